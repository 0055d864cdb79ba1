// MLP building blocks on CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16), shared by the surrogate training kernels
// and the batched material kernels.
//
// Layout ("transposed" formulation): a wavefront evaluates 32 samples at once. A layer computes Z = W H + b with the
// weights as the MFMA A operand (32 output features x 16 inputs per instruction) and the activations as the B operand
// (16 inputs x 32 samples). Measured operand layouts on gfx950 (lane l, element i; h = l / 32):
//   A[m = l % 32][k = 8 h + i],  B[k = 8 h + i][n = l % 32],  D register j: [m = 8 (j / 4) + 4 h + j % 4][n = l % 32].
// A 32-feature output tile in D layout is, after bias + GELU + conversion to BF16, already the B operand of the next
// layer for two 16-input blocks, provided the next layer's weight columns are stored in the matching order: input block
// kb = 2 t + s, lane half h, element i holds hidden feature
//   perm(kb, h, i) = 32 t + 16 s + 8 (i / 4) + 4 h + i % 4.
// So a whole network runs in registers, with no LDS traffic between layers. The backward pass chains the same way with
// W^T packed in the same permuted order. The weight-gradient GEMM (sum over samples) uses the same instruction with
// samples as the reduction dimension.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fecx.h"

namespace fecx {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// Network shapes. D: true widths; P: padded widths (inputs to a 16-multiple, outputs to a 32-multiple).
template <int ARCH>
struct Arch;
template <>
struct Arch<FECX_MLP_J2> {
  static constexpr int NL = 4;
  static constexpr int D[5] = {13, 128, 128, 128, 13};
  static constexpr int P[5] = {16, 128, 128, 128, 32};
};
template <>
struct Arch<FECX_MLP_NH> {
  static constexpr int NL = 3;
  static constexpr int D[4] = {9, 256, 256, 15};
  static constexpr int P[4] = {16, 256, 256, 32};
};

// widest hidden layer (register arrays are sized by it)
template <class A>
constexpr int max_hidden() {
  int w = 32;
  for (int l = 1; l < A::NL; ++l) w = A::P[l] > w ? A::P[l] : w;
  return w;
}

template <class A>
struct Offs {
  // packed forward weights of layer l, in bf16x8 units: (P[l+1] / 32) x (P[l] / 16) x 64 lanes
  static constexpr int64_t wf(int l) {
    int64_t o = 0;
    for (int k = 0; k < l; ++k) o += (int64_t)(A::P[k + 1] / 32) * (A::P[k] / 16) * 64;
    return o;
  }
  // packed backward weights W_l^T (l >= 1): (P[l] / 32) x (P[l+1] / 16) x 64 lanes
  static constexpr int64_t wb(int l) {
    int64_t o = 0;
    for (int k = 1; k < l; ++k) o += (int64_t)(A::P[k] / 32) * (A::P[k + 1] / 16) * 64;
    return o;
  }
  static constexpr int bias(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += A::P[k + 1];
    return o;
  }
  // canonical parameters: per layer W (D[l+1] x D[l], row-major) then b (D[l+1])
  static constexpr int64_t param(int l) {
    int64_t o = 0;
    for (int k = 0; k < l; ++k) o += (int64_t)A::D[k + 1] * A::D[k] + A::D[k + 1];
    return o;
  }
  static constexpr int64_t wf_total() { return wf(A::NL); }
  static constexpr int64_t wb_total() { return wb(A::NL); }
  static constexpr int bias_total() { return bias(A::NL); }
  static constexpr int64_t params() { return param(A::NL); }
  // activation rows stored for the weight gradient: h_l for l < NL (h_0 padded to 32 rows), dz_l for 1 <= l <= NL
  static constexpr int hrows(int l) { return l == 0 ? 32 : A::P[l]; }
};

// hidden-feature order of a packed B operand (see the header comment)
__host__ __device__ constexpr int perm_feature(int kb, int h, int i) {
  return 32 * (kb / 2) + 16 * (kb % 2) + 8 * (i / 4) + 4 * h + i % 4;
}
// feature held by D-layout register j of output tile mt in lane half h
__host__ __device__ constexpr int dreg_feature(int mt, int h, int j) { return 32 * mt + 8 * (j / 4) + 4 * h + j % 4; }

struct MlpDev {
  const bf16x8* wf;        // packed forward weights
  const bf16x8* wb;        // packed W^T for the backward pass
  const float* bias;       // padded biases
  const float* in_shift;   // D[0]
  const float* in_scale;
  const float* out_shift;  // D[NL]
  const float* out_scale;
};

// GELU in its tanh form, 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) = z sigma(u), u = 2 sqrt(2/pi) (z + 0.044715 z^3)
// (Hendrycks & Gimpel, "Gaussian Error Linear Units"), evaluated with the hardware exp2 and reciprocal: five VALU
// operations and two transcendental ones per activation (the erf form costs several times that, and the network's
// hidden activations dominate its VALU work). exp2 overflow gives sigma = 0, the correct limit.
constexpr float GELU_U = 1.5957691216057308f;   // 2 sqrt(2 / pi)
constexpr float GELU_C = 0.044715f;
constexpr float LOG2E = 1.4426950408889634f;
// sigma(u) = 1 / (1 + 2^(-log2(e) u)) with the constants folded: -log2(e) u = z (GA + GB z^2)
constexpr float GELU_A = -LOG2E * GELU_U, GELU_B = GELU_A * GELU_C;
__device__ __forceinline__ float gelu_sigma(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(GELU_B, z * z, GELU_A)));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }
__device__ __forceinline__ float gelu_prime(float z) {
  const float s = gelu_sigma(z);
  const float du = GELU_U * fmaf(3.0f * GELU_C * z, z, 1.0f);
  return fmaf(z * s * (1.0f - s), du, s);
}

// Z = W_l H + b for one layer; hb holds P[l] / 16 B operands, acc receives P[l+1] / 32 D tiles.
template <class A, int L>
__device__ __forceinline__ void mlp_layer(const MlpDev& m, const bf16x8* hb, f32x16* acc, int lane) {
  constexpr int IN = A::P[L], OUT = A::P[L + 1];
  const bf16x8* w = m.wf + Offs<A>::wf(L);
  const float* b = m.bias + Offs<A>::bias(L);
  const int h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < OUT / 32; ++mt) {
    f32x16 c;
#pragma unroll
    for (int j = 0; j < 16; ++j) c[j] = b[dreg_feature(mt, h, j)];
#pragma unroll
    for (int kb = 0; kb < IN / 16; ++kb)
      c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[(mt * (IN / 16) + kb) * 64 + lane], hb[kb], c, 0, 0, 0);
    acc[mt] = c;
  }
}

// dH = W_l^T dZ for the backward pass (l >= 1); db holds P[l+1] / 16 B operands of dZ_{l+1}.
template <class A, int L>
__device__ __forceinline__ void mlp_layer_t(const MlpDev& m, const bf16x8* db, f32x16* acc, int lane) {
  constexpr int IN = A::P[L + 1], OUT = A::P[L];
  const bf16x8* w = m.wb + Offs<A>::wb(L);
#pragma unroll
  for (int mt = 0; mt < OUT / 32; ++mt) {
    f32x16 c = {};
#pragma unroll
    for (int kb = 0; kb < IN / 16; ++kb)
      c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[(mt * (IN / 16) + kb) * 64 + lane], db[kb], c, 0, 0, 0);
    acc[mt] = c;
  }
}

// D tiles -> B operands of the next layer (no activation)
template <int OUT>
__device__ __forceinline__ void pack_tiles(const f32x16* acc, bf16x8* hb) {
#pragma unroll
  for (int mt = 0; mt < OUT / 32; ++mt)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) hb[2 * mt + s][i] = (__bf16)acc[mt][8 * s + i];
}

// GELU in place on the D tiles, then pack as B operands
template <int OUT>
__device__ __forceinline__ void gelu_pack(const f32x16* acc, bf16x8* hb) {
#pragma unroll
  for (int mt = 0; mt < OUT / 32; ++mt)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) hb[2 * mt + s][i] = (__bf16)gelu(acc[mt][8 * s + i]);
}

// Z = W_l H + b for one hidden layer, with the bias, GELU and BF16 packing applied one 32-feature D tile at a time
// straight into the next layer's B operands: one tile (16 registers) is live instead of P[l+1] / 2, so the inference
// kernels fit more wavefronts per SIMD and one wavefront's activations overlap another's MFMAs. Same arithmetic, in the
// same order, as mlp_layer followed by gelu_pack.
template <class A, int L>
__device__ __forceinline__ void mlp_layer_gelu(const MlpDev& m, const bf16x8* hb, bf16x8* hn, int lane) {
  constexpr int IN = A::P[L], OUT = A::P[L + 1];
  const bf16x8* w = m.wf + Offs<A>::wf(L);
  const float* b = m.bias + Offs<A>::bias(L);
  const int h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < OUT / 32; ++mt) {
    f32x16 c;
#pragma unroll
    for (int j = 0; j < 16; ++j) c[j] = b[dreg_feature(mt, h, j)];
#pragma unroll
    for (int kb = 0; kb < IN / 16; ++kb)
      c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[(mt * (IN / 16) + kb) * 64 + lane], hb[kb], c, 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) hn[2 * mt + s][i] = (__bf16)gelu(c[8 * s + i]);
  }
}

// Forward pass of one wavefront (32 samples). xin: this lane's sample features, already normalized (the first D[0]
// of them); on return y holds all P[NL] normalized outputs of this lane's sample (both lane halves agree).
template <class A>
__device__ __forceinline__ void mlp_forward_wave(const MlpDev& m, const float* xin, float* y, int lane) {
  static_assert(A::NL == 3 || A::NL == 4, "three or four layers");
  static_assert(A::P[0] == 16 && A::P[A::NL] == 32, "one input block and one output tile");
  const int h = lane >> 5;
  bf16x8 h0[1];
#pragma unroll
  for (int i = 0; i < 8; ++i) {  // lane half h takes features 8 h .. 8 h + 7 (compile-time indices only)
    const float v0 = i < A::D[0] ? xin[i] : 0.0f;
    const float v1 = 8 + i < A::D[0] ? xin[(8 + i) % A::D[0]] : 0.0f;
    h0[0][i] = (__bf16)(h ? v1 : v0);
  }
  bf16x8 h1[A::P[1] / 16];
  mlp_layer_gelu<A, 0>(m, h0, h1, lane);
  f32x16 out[1];
  if constexpr (A::NL == 3) {
    bf16x8 h2[A::P[2] / 16];
    mlp_layer_gelu<A, 1>(m, h1, h2, lane);
    mlp_layer<A, 2>(m, h2, out, lane);
  } else {
    bf16x8 h2[A::P[2] / 16], h3[A::P[3] / 16];
    mlp_layer_gelu<A, 1>(m, h1, h2, lane);
    mlp_layer_gelu<A, 2>(m, h2, h3, lane);
    mlp_layer<A, 3>(m, h3, out, lane);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const float mine = out[0][j], other = __shfl_xor(mine, 32);
    y[dreg_feature(0, 0, j)] = h ? other : mine;
    y[dreg_feature(0, 1, j)] = h ? mine : other;
  }
}

}  // namespace fecx
