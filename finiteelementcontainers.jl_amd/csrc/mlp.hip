// Learned constitutive surrogate: training and batched inference kernels (BF16 MFMA, FP32 accumulation), the batched
// plain material updates they are trained on and fall back to, and training-sample generation.
//
// The reference has no surrogate and no batched material pass; its materials are evaluated inside the element loop
// (test/mechanics/TestMechanicsCommon.jl:140-222). The plain models here are the same FP32 material functions the
// element kernels use (fem_device.h), so a surrogate is checked against exactly what the plain assembly computes.
#include "fem_device.h"
#include "mlp_device.h"
#include "surrogate.h"

namespace fecx {

// ------------------------------------------------------------------ per-material feature maps
// J2: raw input {H (9, row-major displacement gradient), state (7)}, features {eps (6), plastic strain (6), eqps},
//     raw output {sigma (6), plastic strain increment (6), eqps increment}; symmetric tensors in the state order
//     (11, 21, 31, 22, 32, 33).
// NH: raw input H (9), features H, raw output {P (9, row-major), C^-1 - I (6)}.
template <int ARCH>
struct Feat;
template <>
struct Feat<FECX_MLP_J2> {
  static constexpr int IN_W = 16, NF = 13, OUT_W = 13;
  __device__ static void features(const float* raw, float* x) {
    x[0] = raw[0]; x[1] = 0.5f * (raw[3] + raw[1]); x[2] = 0.5f * (raw[6] + raw[2]);
    x[3] = raw[4]; x[4] = 0.5f * (raw[7] + raw[5]); x[5] = raw[8];
#pragma unroll
    for (int i = 0; i < 7; ++i) x[6 + i] = raw[9 + i];
  }
};
template <>
struct Feat<FECX_MLP_NH> {
  static constexpr int IN_W = 9, NF = 9, OUT_W = 15;
  __device__ static void features(const float* raw, float* x) {
#pragma unroll
    for (int i = 0; i < 9; ++i) x[i] = raw[i];
  }
};

template <int ARCH>
__device__ __forceinline__ void normalized_features(const MlpDev& m, const float* raw, float* x) {
  Feat<ARCH>::features(raw, x);
#pragma unroll
  for (int f = 0; f < Feat<ARCH>::NF; ++f) x[f] = (x[f] - m.in_shift[f]) / m.in_scale[f];
}

// ------------------------------------------------------------------ training: fused forward + backward per wavefront
template <int ARCH>
__global__ __launch_bounds__(256) void mlp_train_kernel(TrainArgs a) {
  using A = Arch<ARCH>;
  using O = Offs<A>;
  constexpr int NL = A::NL, DOUT = A::D[NL];
  constexpr int MAXW = max_hidden<A>();
  const int lane = threadIdx.x % 64, h = lane >> 5;
  const int64_t tile = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;  // 32 samples per wavefront
  const int64_t s = tile * 32 + lane % 32;
  if (tile * 32 >= a.npad) return;  // whole wavefront out of range (npad is a multiple of 128: never taken)
  const bool valid = s < a.n;
  const int64_t row = valid ? (a.idx ? (int64_t)a.idx[s] : s) : 0;
  float raw[Feat<ARCH>::IN_W], x[16];
#pragma unroll
  for (int i = 0; i < Feat<ARCH>::IN_W; ++i) raw[i] = valid ? a.X[row * Feat<ARCH>::IN_W + i] : 0.0f;
  normalized_features<ARCH>(a.m, raw, x);
  if (!valid) {
#pragma unroll
    for (int f = 0; f < 16; ++f) x[f] = 0.0f;
  }
  bf16x8 hb[MAXW / 16];
  f32x16 acc[MAXW / 32];
  // h_0: rows 8h .. 8h+7 from this lane half, rows 16 + 8h .. 23 + 8h zero
  __bf16* h0 = a.hbuf;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float v0 = i < A::D[0] ? x[i] : 0.0f;
    const float v1 = 8 + i < A::D[0] ? x[(8 + i) % A::D[0]] : 0.0f;
    hb[0][i] = (__bf16)(h ? v1 : v0);
    h0[(int64_t)(8 * h + i) * a.npad + s] = hb[0][i];
    h0[(int64_t)(16 + 8 * h + i) * a.npad + s] = (__bf16)0.0f;
  }
  // forward, keeping GELU'(z) per hidden layer
  int64_t hoff = (int64_t)O::hrows(0) * a.npad, gpoff = 0;
  auto hidden = [&](auto LC) {
    constexpr int L = decltype(LC)::value;  // hidden layer index 1 .. NL-1
    constexpr int W = A::P[L];
    __bf16* hl = a.hbuf + hoff;
    __bf16* gp = a.gpbuf + gpoff + tile * (W / 2) * 64;  // [tile][register r = 16 mt + j][lane]
#pragma unroll
    for (int mt = 0; mt < W / 32; ++mt)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float z = acc[mt][j];
        const float g = gelu(z);
        gp[(mt * 16 + j) * 64 + lane] = (__bf16)gelu_prime(z);
        hl[(int64_t)dreg_feature(mt, h, j) * a.npad + s] = (__bf16)g;
        hb[2 * mt + j / 8][j % 8] = (__bf16)g;
      }
    hoff += (int64_t)W * a.npad;
    gpoff += (int64_t)(a.npad / 32) * (W / 2) * 64;
  };
  mlp_layer<A, 0>(a.m, hb, acc, lane);
  hidden(std::integral_constant<int, 1>());
  mlp_layer<A, 1>(a.m, hb, acc, lane);
  if constexpr (NL >= 3) {
    hidden(std::integral_constant<int, 2>());
    mlp_layer<A, 2>(a.m, hb, acc, lane);
  }
  if constexpr (NL >= 4) {
    hidden(std::integral_constant<int, 3>());
    mlp_layer<A, 3>(a.m, hb, acc, lane);
  }
  // loss and output gradient, dz_NL = 2 (y - t) / (n * DOUT)
  const float gscale = 2.0f / (float)((double)a.n * DOUT);
  double lsum = 0.0;
  int64_t dzoff = 0;
  for (int l = 1; l < NL; ++l) dzoff += (int64_t)A::P[l] * a.npad;
  __bf16* dzl = a.dzbuf + dzoff;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int o = dreg_feature(0, h, j);
    float dy = 0.0f;
    if (valid && o < DOUT) {
      const float t = (a.T[row * DOUT + o] - a.m.out_shift[o]) / a.m.out_scale[o];
      const float d = acc[0][j] - t;
      lsum += (double)d * d;
      dy = gscale * d;
    }
    acc[0][j] = dy;
    dzl[(int64_t)o * a.npad + s] = (__bf16)dy;
  }
  // wave sum of the loss, in a fixed order
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0) a.loss_part[tile] = lsum / ((double)a.n * DOUT);
  // backward
  pack_tiles<32>(acc, hb);
  auto back = [&](auto LC) {
    constexpr int L = decltype(LC)::value;  // dh_L = W_L^T dz_{L+1}; dz_L = dh_L * GELU'(z_L)
    constexpr int W = A::P[L];
    mlp_layer_t<A, L>(a.m, hb, acc, lane);
    gpoff -= (int64_t)(a.npad / 32) * (W / 2) * 64;
    dzoff -= (int64_t)W * a.npad;
    const __bf16* gp = a.gpbuf + gpoff + tile * (W / 2) * 64;
    __bf16* dz = a.dzbuf + dzoff;
#pragma unroll
    for (int mt = 0; mt < W / 32; ++mt)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float d = acc[mt][j] * (float)gp[(mt * 16 + j) * 64 + lane];
        dz[(int64_t)dreg_feature(mt, h, j) * a.npad + s] = (__bf16)d;
        hb[2 * mt + j / 8][j % 8] = (__bf16)d;
      }
  };
  if constexpr (NL >= 4) back(std::integral_constant<int, 3>());
  if constexpr (NL >= 3) back(std::integral_constant<int, 2>());
  back(std::integral_constant<int, 1>());
}

// ------------------------------------------------------------------ training: weight gradients (MFMA over samples)
// One wavefront per (layer, 32 x 32 weight tile, chunk of samples): dW[o][i] = sum_s dz_{l+1}[o][s] h_l[i][s] with
// samples as the MFMA reduction dimension; the tiles with it == 0 also sum dz rows for the bias gradient. Partial sums
// per chunk go to part[chunk][param]; the reduction over chunks runs in a fixed order (mlp_update_kernel).
template <int ARCH>
__global__ __launch_bounds__(64) void mlp_wgrad_kernel(WGradArgs a) {
  using A = Arch<ARCH>;
  using O = Offs<A>;
  constexpr int NL = A::NL;
  const int lane = threadIdx.x, h = lane >> 5;
  int t = blockIdx.x;
  const int chunk = blockIdx.y;
  const int64_t s0 = (int64_t)chunk * a.chunk_len;
  int64_t hoff = 0, dzoff = 0;
  auto tile = [&](auto LC) {
    constexpr int L = decltype(LC)::value;
    constexpr int NOT = A::P[L + 1] / 32, NIT = O::hrows(L) / 32;
    if (t < 0) return;
    if (t >= NOT * NIT) {
      t -= NOT * NIT;
      hoff += (int64_t)O::hrows(L) * a.npad;
      dzoff += (int64_t)A::P[L + 1] * a.npad;
      return;
    }
    const int ot = t / NIT, it = t % NIT;
    t = -1;
    const __bf16* dz = a.dzbuf + dzoff + (int64_t)(ot * 32 + lane % 32) * a.npad;
    const __bf16* hl = a.hbuf + hoff + (int64_t)(it * 32 + lane % 32) * a.npad;
    f32x16 c = {};
    float bsum = 0.0f;
    for (int64_t k = s0; k < s0 + a.chunk_len; k += 16) {
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(dz + k + 8 * h);
      const bf16x8 bv = *reinterpret_cast<const bf16x8*>(hl + k + 8 * h);
      c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
      if (it == 0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) bsum += (float)av[e];
      }
    }
    float* part = a.part + (int64_t)chunk * O::params() + O::param(L);
    const int i = it * 32 + lane % 32;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int o = dreg_feature(ot, h, j);
      if (o < A::D[L + 1] && i < A::D[L]) part[(int64_t)o * A::D[L] + i] = c[j];
    }
    if (it == 0) {
      bsum += __shfl_xor(bsum, 32);
      const int o = ot * 32 + lane % 32;
      if (h == 0 && o < A::D[L + 1]) part[(int64_t)A::D[L + 1] * A::D[L] + o] = bsum;
    }
  };
  tile(std::integral_constant<int, 0>());
  tile(std::integral_constant<int, 1>());
  if constexpr (NL >= 3) tile(std::integral_constant<int, 2>());
  if constexpr (NL >= 4) tile(std::integral_constant<int, 3>());
}

template <int ARCH>
int wgrad_tiles() {
  using A = Arch<ARCH>;
  int n = 0;
  for (int l = 0; l < A::NL; ++l) n += (A::P[l + 1] / 32) * (Offs<A>::hrows(l) / 32);
  return n;
}

// ------------------------------------------------------------------ parameter update and repacking
// Writes canonical parameter p into the packed forward / backward weights or the padded biases.
template <int ARCH>
__device__ void repack_param(const MlpPack& pk, int64_t p, float w) {
  using A = Arch<ARCH>;
  using O = Offs<A>;
  int L = 0;
  while (L + 1 < A::NL && p >= O::param(L + 1)) ++L;
  int64_t q = p - O::param(L);
  const int din = A::D[L], dout = A::D[L + 1];
  __bf16* wf = reinterpret_cast<__bf16*>(pk.wf);
  __bf16* wb = reinterpret_cast<__bf16*>(pk.wb);
  if (q < (int64_t)dout * din) {
    const int o = (int)(q / din), i = (int)(q % din);
    // forward: A operand row o, input i
    int kb, hh, e;
    if (L == 0) { kb = i / 16; hh = (i % 16) / 8; e = i % 8; }
    else { const int t = i / 32, r = i % 32, sgn = r / 16, rem = r % 16; kb = 2 * t + sgn; e = 4 * (rem / 8) + rem % 4; hh = (rem % 8) / 4; }
    const int64_t vf = O::wf(L) + ((int64_t)(o / 32) * (A::P[L] / 16) + kb) * 64 + hh * 32 + o % 32;
    wf[vf * 8 + e] = (__bf16)w;
    if (L >= 1) {  // backward: W^T row i, input o (permuted order of layer L's outputs)
      const int t = o / 32, r = o % 32, sgn = r / 16, rem = r % 16;
      const int kb2 = 2 * t + sgn, e2 = 4 * (rem / 8) + rem % 4, h2 = (rem % 8) / 4;
      const int64_t vb = O::wb(L) + ((int64_t)(i / 32) * (A::P[L + 1] / 16) + kb2) * 64 + h2 * 32 + i % 32;
      wb[vb * 8 + e2] = (__bf16)w;
    }
  } else {
    pk.bias[O::bias(L) + (q - (int64_t)dout * din)] = w;
  }
}

template <int ARCH>
__global__ void mlp_update_kernel(UpdateArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  float w = a.w[p];
  if (a.part) {
    float g = 0.0f;
    for (int c = 0; c < a.nchunk; ++c) g += a.part[(int64_t)c * a.np + p];
    if (a.grad) a.grad[p] = g;
    if (a.adam) {
      const float m1 = a.b1 * a.m1[p] + (1.0f - a.b1) * g;
      const float m2 = a.b2 * a.m2[p] + (1.0f - a.b2) * g * g;
      a.m1[p] = m1;
      a.m2[p] = m2;
      w -= a.lr * (m1 * a.c1) / (sqrtf(m2 * a.c2) + a.eps);
      a.w[p] = w;
    }
  }
  if (a.repack) repack_param<ARCH>(a.pk, p, w);
}

// Sum of the n loss partials by one wavefront in a fixed order (lane l takes partials l, l + 64, ..., then a fixed
// xor tree across lanes), so every run gives the same bits.
__global__ void loss_sum_kernel(const double* part, int64_t n, double* out) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 64) s += part[i];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x == 0) out[0] = s;
}

__global__ void gather_rows_kernel(int64_t n, int w, const int32_t* idx, const float* src, float* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * w) dst[i] = src[(int64_t)idx[i / w] * w + i % w];
}

// ------------------------------------------------------------------ plain material models on one sample
// J2 raw in {H (9), state (7)} -> raw out {sigma (6), dplastic (6), deqps}, plus P (9), state_new (7), C (9)
__device__ __forceinline__ void j2_plain(const MatParams& mp, const float* H, const float* so, float* y, float* P,
                                         float* sn, float* C) {
  j2_update(mp.p, H, so, sn, P, C);
  y[0] = P[0]; y[1] = P[3]; y[2] = P[6]; y[3] = P[4]; y[4] = P[7]; y[5] = P[8];
#pragma unroll
  for (int i = 0; i < 7; ++i) y[6 + i] = sn[i] - so[i];
}

// NH raw in H (9) -> raw out {P (9), C^-1 - I (6)}, plus C = {F^-T (9), mu - lambda ln J}
__device__ __forceinline__ void nh_plain(const MatParams& mp, const float* H, float* y, float* P, float* C, bool& bad) {
  Mat<FECX_NEOHOOKEAN>::eval<3>(mp, H, nullptr, nullptr, P, C, bad);
#pragma unroll
  for (int i = 0; i < 9; ++i) y[i] = P[i];
  // C^-1 = F^-1 F^-T, F^-T row-major in C[0..9)
  float Ci[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Ci[i * 3 + j] = C[0 * 3 + i] * C[0 * 3 + j] + C[1 * 3 + i] * C[1 * 3 + j] + C[2 * 3 + i] * C[2 * 3 + j];
  y[9] = Ci[0] - 1.0f; y[10] = Ci[3]; y[11] = Ci[6]; y[12] = Ci[4] - 1.0f; y[13] = Ci[7]; y[14] = Ci[8] - 1.0f;
}

// ------------------------------------------------------------------ surrogate decode + residual check
// J2: the network's {sigma, dplastic, deqps} is accepted when it satisfies the return-mapping equations to `tol`:
//   elastic law   |sigma - (kappa tr(eps - ep_new) I + 2 mu dev(eps - ep_new))| <= tol sigma_y
//   yield         |f| <= tol sigma_y when deqps > 0 (on the yield surface), f <= tol sigma_y otherwise
//   flow          |dplastic - sqrt(3/2) deqps n| <= tol sigma_y / (2 mu), deqps >= -tol sigma_y / (2 mu)
// The tangent data follow from the closed-form consistent tangent at the network's plastic multiplier.
__device__ __forceinline__ bool j2_surrogate(const MatParams& mp, const float* H, const float* so, const float* y,
                                             float tol, float* P, float* sn, float* C) {
  const float E = mp.p[0], nu = mp.p[1], sy = mp.p[2], Hh = mp.p[3];
  const float mu = E / (2 * (1 + nu)), kap = E / (3 * (1 - 2 * nu));
  const float r23 = 0.81649658092772603f, r32 = 1.2247448713915890f;
  const float eref = sy / (2 * mu);
  auto sym = [](const float* v, float* M) {  // state order -> row-major
    M[0] = v[0]; M[1] = v[1]; M[2] = v[2]; M[3] = v[1]; M[4] = v[3]; M[5] = v[4]; M[6] = v[2]; M[7] = v[4]; M[8] = v[5];
  };
  float sig[9], dep[9], epo[9], eps[9];
  sym(y, sig);
  sym(y + 6, dep);
  sym(so, epo);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) eps[i * 3 + j] = 0.5f * (H[i * 3 + j] + H[j * 3 + i]);
  const float dal = y[12];
  // elastic law
  float ee[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) ee[i] = eps[i] - epo[i] - dep[i];
  const float tre = ee[0] + ee[4] + ee[8];
  float r = 0.0f;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float se = 2 * mu * ee[i] + ((i % 4 == 0) ? (kap - 2 * mu / 3) * tre : 0.0f);
    r = fmaxf(r, fabsf(sig[i] - se) / sy);
  }
  // yield
  const float trs = (sig[0] + sig[4] + sig[8]) / 3;
  float s[9], sn2 = 0.0f;
#pragma unroll
  for (int i = 0; i < 9; ++i) { s[i] = sig[i] - ((i % 4 == 0) ? trs : 0.0f); sn2 += s[i] * s[i]; }
  const float snorm = sqrtf(sn2);
  const float f = snorm - r23 * (sy + Hh * (so[6] + dal));
  const bool plastic = dal > tol * eref;
  r = fmaxf(r, plastic ? fabsf(f) / sy : fmaxf(f, 0.0f) / sy);
  // flow rule
  const float inv = snorm > 0.0f ? 1.0f / snorm : 0.0f;
#pragma unroll
  for (int i = 0; i < 9; ++i) r = fmaxf(r, fabsf(dep[i] - r32 * fmaxf(dal, 0.0f) * s[i] * inv) / eref);
  r = fmaxf(r, fmaxf(-dal, 0.0f) / eref);
  if (!(r <= tol)) return false;
#pragma unroll
  for (int i = 0; i < 9; ++i) P[i] = sig[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) sn[i] = so[i] + y[6 + i];
  sn[6] = so[6] + y[12];
  // tangent from the trial state and the network's multiplier
  float st[9], stn2 = 0.0f;
  const float trt = (eps[0] - epo[0]) + (eps[4] - epo[4]) + (eps[8] - epo[8]);
#pragma unroll
  for (int i = 0; i < 9; ++i) { st[i] = 2 * mu * (eps[i] - epo[i] - ((i % 4 == 0) ? trt / 3 : 0.0f)); stn2 += st[i] * st[i]; }
  const float stn = sqrtf(stn2);
  const float dg = plastic ? r32 * dal : 0.0f;
  const float th = (plastic && stn > 0.0f) ? 1.0f - 2 * mu * dg / stn : 1.0f;
  const float thb = plastic ? 1.0f / (1.0f + Hh / (3 * mu)) - (1.0f - th) : 0.0f;
  const float ti = stn > 0.0f ? 1.0f / stn : 0.0f;
  C[0] = st[0] * ti; C[1] = st[4] * ti; C[2] = st[8] * ti; C[3] = st[1] * ti; C[4] = st[5] * ti; C[5] = st[2] * ti;
  C[6] = kap - 2 * mu * th / 3; C[7] = mu * th; C[8] = 2 * mu * thb;
  return true;
}

// NH: the network's {P, C^-1 - I} is accepted when C C^-1 = I to `tol` and P agrees to `tol` with the stress the
// predicted C^-1 implies, P(F, C^-1) = mu (F - F C^-1) + lambda ln J F C^-1 (F^-T = F C^-1), relative to the stress
// scale (mu + lambda) max(|H|, 0.01).
__device__ __forceinline__ bool nh_surrogate(const MatParams& mp, const float* H, const float* y, float tol, float* P,
                                             float* C) {
  const float lam = mp.p[0], mu = mp.p[1];
  const float F[9] = {1 + H[0], H[1], H[2], H[3], 1 + H[4], H[5], H[6], H[7], 1 + H[8]};
  const float Ci[9] = {1 + y[9], y[10], y[11], y[10], 1 + y[12], y[13], y[11], y[13], 1 + y[14]};
  float r = 0.0f, hmax = 0.01f;
#pragma unroll
  for (int i = 0; i < 9; ++i) hmax = fmaxf(hmax, fabsf(H[i]));
  // C C^-1 - I with C = F^T F
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float v = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float Cik = F[0 * 3 + i] * F[0 * 3 + k] + F[1 * 3 + i] * F[1 * 3 + k] + F[2 * 3 + i] * F[2 * 3 + k];
        v += Cik * Ci[k * 3 + j];
      }
      r = fmaxf(r, fabsf(v - (i == j ? 1.0f : 0.0f)));
    }
  const float tr = H[0] + H[4] + H[8];
  const float trH2 = H[0] * H[0] + H[4] * H[4] + H[8] * H[8] + 2.0f * (H[1] * H[3] + H[2] * H[6] + H[5] * H[7]);
  const float detH = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) + H[2] * (H[3] * H[7] - H[4] * H[6]);
  const float Jm1 = tr + 0.5f * (tr * tr - trH2) + detH;
  if (!(Jm1 > -1.0f)) return false;
  const float lnJ = log1pf(Jm1);
  float FiT[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) FiT[i * 3 + j] = F[i * 3 + 0] * Ci[0 * 3 + j] + F[i * 3 + 1] * Ci[1 * 3 + j] + F[i * 3 + 2] * Ci[2 * 3 + j];
  const float ps = (mu + lam) * hmax;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float pp = mu * (F[i] - FiT[i]) + lam * lnJ * FiT[i];
    r = fmaxf(r, fabsf(y[i] - pp) / ps);
  }
  if (!(r <= tol)) return false;
#pragma unroll
  for (int i = 0; i < 9; ++i) { P[i] = y[i]; C[i] = FiT[i]; }
  C[9] = mu - lam * lnJ;
  return true;
}

// ------------------------------------------------------------------ batched material update
// One wavefront per 32 samples (both lane halves carry the same sample; half 0 stores). Sources and sinks:
//   batch mode    : in (n x IN_W raw) -> out (n x OUT_W raw)
//   assembly mode : hq (n x 9 quadrature-point gradients, n = ne x nq in element storage order) and the element's
//                   state (via eperm) -> qpmat (n x (9 + NC): P, compact tangent) and state_new
template <int ARCH, bool SURR>
__global__ __launch_bounds__(256) void material_kernel(MatBatchArgs a) {
  using A = Arch<ARCH>;
  using FT = Feat<ARCH>;
  constexpr bool J2 = ARCH == FECX_MLP_J2;
  constexpr int NC = J2 ? Mat<FECX_J2>::NC : Mat<FECX_NEOHOOKEAN>::NC;
  const int lane = threadIdx.x % 64, h = lane >> 5;
  const int64_t tile = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  const int64_t i = tile * 32 + lane % 32;
  if (tile * 32 >= a.n) return;  // wave-uniform
  const bool valid = i < a.n;
  float H[9] = {}, so[7] = {};
  int64_t sidx = 0;
  if (valid) {
    if (a.hq) {
#pragma unroll
      for (int c = 0; c < 9; ++c) H[c] = a.hq[i * 9 + c];
      if (J2) {
        sidx = ((int64_t)a.eperm[i / a.nq] * a.nq + i % a.nq) * 7;
#pragma unroll
        for (int c = 0; c < 7; ++c) so[c] = a.st_old[sidx + c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 9; ++c) H[c] = a.in[i * FT::IN_W + c];
      if (J2) {
#pragma unroll
        for (int c = 0; c < 7; ++c) so[c] = a.in[i * FT::IN_W + 9 + c];
      }
    }
  }
  float P[9], C[10], sn[7], y[16];
  bool ok = false, bad = false;
  if constexpr (SURR) {
    float raw[16], x[16];
#pragma unroll
    for (int c = 0; c < 9; ++c) raw[c] = H[c];
#pragma unroll
    for (int c = 0; c < 7; ++c) raw[9 + c] = so[c];
    normalized_features<ARCH>(a.m, raw, x);
    float yn[32];
    mlp_forward_wave<A>(a.m, x, yn, lane);
#pragma unroll
    for (int o = 0; o < FT::OUT_W; ++o) y[o] = yn[o] * a.m.out_scale[o] + a.m.out_shift[o];
    if (valid) {
      // frozen decisions: a point that passed keeps the network's output (checked with an infinite tolerance, which
      // only assembles P and the tangent data), a point that fell back stays on the plain model
      const bool frozen = a.freeze && a.accept;
      const bool use = frozen ? a.accept[i] != 0 : true;
      const float tol = frozen ? 3.0e38f : a.tol;
      if (use) {
        if constexpr (J2) ok = j2_surrogate(a.mp, H, so, y, tol, P, sn, C);
        else ok = nh_surrogate(a.mp, H, y, tol, P, C);
      }
      if (a.accept && !frozen && h == 0) a.accept[i] = ok ? 1 : 0;
    }
    const uint64_t fb = __ballot(valid && h == 0 && !ok);
    if (lane == 0 && fb && a.nfallback) atomicAdd(a.nfallback, (unsigned long long)__popcll(fb));
  }
  if (valid && !ok) {  // the plain model: the fallback, or the whole update when SURR is off
    if constexpr (J2) j2_plain(a.mp, H, so, y, P, sn, C);
    else nh_plain(a.mp, H, y, P, C, bad);
  } else if (valid && SURR && !a.hq) {
    // batch output of an accepted surrogate sample: report what the assembly would use
    if constexpr (J2) {
      y[0] = P[0]; y[1] = P[3]; y[2] = P[6]; y[3] = P[4]; y[4] = P[7]; y[5] = P[8];
#pragma unroll
      for (int c = 0; c < 7; ++c) y[6 + c] = sn[c] - so[c];
    }
  }
  if (!valid || h != 0) return;
  if (bad && a.flag) atomicOr(a.flag, 1);
  if (a.hq) {
    float* q = a.qpmat + i * (9 + NC);
#pragma unroll
    for (int c = 0; c < 9; ++c) q[c] = P[c];
#pragma unroll
    for (int c = 0; c < NC; ++c) q[9 + c] = C[c];
    if (J2 && a.st_new) {
#pragma unroll
      for (int c = 0; c < 7; ++c) a.st_new[sidx + c] = sn[c];
    }
  } else {
#pragma unroll
    for (int c = 0; c < FT::OUT_W; ++c) a.out[i * FT::OUT_W + c] = y[c];
  }
}

// ------------------------------------------------------------------ network-only forward (raw units)
template <int ARCH>
__global__ __launch_bounds__(256) void mlp_forward_kernel(MatBatchArgs a) {
  using A = Arch<ARCH>;
  using FT = Feat<ARCH>;
  const int lane = threadIdx.x % 64, h = lane >> 5;
  const int64_t tile = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  const int64_t i = tile * 32 + lane % 32;
  if (tile * 32 >= a.n) return;
  const bool valid = i < a.n;
  float raw[16] = {}, x[16], yn[32];
#pragma unroll
  for (int c = 0; c < FT::IN_W; ++c) raw[c] = valid ? a.in[i * FT::IN_W + c] : 0.0f;
  normalized_features<ARCH>(a.m, raw, x);
  mlp_forward_wave<A>(a.m, x, yn, lane);
  if (valid && h == 0) {
#pragma unroll
    for (int o = 0; o < FT::OUT_W; ++o) a.out[i * FT::OUT_W + o] = yn[o] * a.m.out_scale[o] + a.m.out_shift[o];
  }
}

// ------------------------------------------------------------------ J2 training samples: random strain paths
// Thread p follows one macroscopic strain path from zero through 4 random waypoints (symmetric directions, norm
// max_norm * U(0.2, 1.1)), piecewise linear over `steps` steps, updating the plain model; every step emits the raw
// input {H = eps, state} and the raw output {sigma, dplastic, deqps}.
__device__ __forceinline__ uint64_t smix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float urand(uint64_t& s) { return (float)((smix(s) >> 40) * (1.0 / 16777216.0)); }
__device__ __forceinline__ float nrand(uint64_t& s) {
  const float u1 = fmaxf(urand(s), 1e-7f), u2 = urand(s);
  return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

__global__ void j2_paths_kernel(int64_t npaths, int steps, float max_norm, uint64_t seed, MatParams mp, float* X,
                                float* T) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npaths) return;
  uint64_t st = seed ^ ((uint64_t)p * 0xD1B54A32D192ED03ULL);
  float W[5][9];
#pragma unroll
  for (int c = 0; c < 9; ++c) W[0][c] = 0.0f;
  for (int k = 1; k < 5; ++k) {
    float g[6], nn = 0.0f;
    for (int c = 0; c < 6; ++c) g[c] = nrand(st);
    float S[9] = {g[0], g[3], g[5], g[3], g[1], g[4], g[5], g[4], g[2]};
    for (int c = 0; c < 9; ++c) nn += S[c] * S[c];
    const float amp = max_norm * (0.2f + 0.9f * urand(st)) / sqrtf(nn);
    for (int c = 0; c < 9; ++c) W[k][c] = amp * S[c];
  }
  const int per = steps / 4 > 0 ? steps / 4 : 1;
  float so[7] = {}, sn[7], P[9], C[9], y[13];
  for (int n = 1; n <= steps; ++n) {
    const int seg = min((n - 1) / per, 3);
    const float t = (float)(n - seg * per) / (float)per;
    float Hs[9];
    for (int c = 0; c < 9; ++c) Hs[c] = (1 - t) * W[seg][c] + t * W[seg + 1][c];
    j2_plain(mp, Hs, so, y, P, sn, C);
    const int64_t r = p * steps + (n - 1);
    for (int c = 0; c < 9; ++c) X[r * 16 + c] = Hs[c];
    for (int c = 0; c < 7; ++c) X[r * 16 + 9 + c] = so[c];
    for (int c = 0; c < 13; ++c) T[r * 13 + c] = y[c];
    for (int c = 0; c < 7; ++c) so[c] = sn[c];
  }
}

// ------------------------------------------------------------------ column mean / standard deviation (FP64 sums)
__global__ void column_stats_kernel(const float* X, int64_t n, int w, int ncol, float* shift, float* scale) {
  const int c = blockIdx.x;  // one work-group per column
  if (c >= ncol) return;
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = X[i * w + c];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) { s1[threadIdx.x] += s1[threadIdx.x + off]; s2[threadIdx.x] += s2[threadIdx.x + off]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = s1[0] / (double)n, var = fmax(s2[0] / (double)n - m * m, 0.0);
    const double sd = sqrt(var);
    shift[c] = (float)m;
    scale[c] = (float)(sd > 1e-12 * fmax(1.0, fabs(m)) ? sd : 1.0);
  }
}

// J2 features from raw inputs, for the column statistics of the inputs
__global__ void j2_feature_rows_kernel(const float* in, int64_t n, float* x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float raw[16], f[13];
  for (int c = 0; c < 16; ++c) raw[c] = in[i * 16 + c];
  Feat<FECX_MLP_J2>::features(raw, f);
  for (int c = 0; c < 13; ++c) x[i * 13 + c] = f[c];
}

// ------------------------------------------------------------------ host launchers
template <int ARCH>
static void launch_train(const TrainArgs& a, hipStream_t s) {
  mlp_train_kernel<ARCH><<<(unsigned)(a.npad / 128), 256, 0, s>>>(a);
}
template <int ARCH>
static void launch_wgrad(const WGradArgs& a, hipStream_t s) {
  dim3 g((unsigned)wgrad_tiles<ARCH>(), (unsigned)a.nchunk);
  mlp_wgrad_kernel<ARCH><<<g, 64, 0, s>>>(a);
}

void mlp_launch_train(int arch, const TrainArgs& a, hipStream_t s) {
  if (arch == FECX_MLP_J2) launch_train<FECX_MLP_J2>(a, s);
  else launch_train<FECX_MLP_NH>(a, s);
}
void mlp_launch_wgrad(int arch, const WGradArgs& a, hipStream_t s) {
  if (arch == FECX_MLP_J2) launch_wgrad<FECX_MLP_J2>(a, s);
  else launch_wgrad<FECX_MLP_NH>(a, s);
}
void mlp_launch_update(int arch, const UpdateArgs& a, hipStream_t s) {
  const unsigned g = (unsigned)((a.np + 255) / 256);
  if (arch == FECX_MLP_J2) mlp_update_kernel<FECX_MLP_J2><<<g, 256, 0, s>>>(a);
  else mlp_update_kernel<FECX_MLP_NH><<<g, 256, 0, s>>>(a);
}
void mlp_launch_loss_sum(const double* part, int64_t n, double* out, hipStream_t s) {
  loss_sum_kernel<<<1, 64, 0, s>>>(part, n, out);
}
void mlp_launch_gather(int64_t n, int w, const int32_t* idx, const float* src, float* dst, hipStream_t s) {
  if (n > 0) gather_rows_kernel<<<(unsigned)((n * w + 255) / 256), 256, 0, s>>>(n, w, idx, src, dst);
}
void mlp_launch_forward(int arch, const MatBatchArgs& a, hipStream_t s) {
  const unsigned g = (unsigned)((a.n + 127) / 128);
  if (a.n == 0) return;
  if (arch == FECX_MLP_J2) mlp_forward_kernel<FECX_MLP_J2><<<g, 256, 0, s>>>(a);
  else mlp_forward_kernel<FECX_MLP_NH><<<g, 256, 0, s>>>(a);
}
void mlp_launch_material(int arch, bool surrogate, const MatBatchArgs& a, hipStream_t s) {
  const unsigned g = (unsigned)((a.n + 127) / 128);
  if (a.n == 0) return;
  if (arch == FECX_MLP_J2) {
    if (surrogate) material_kernel<FECX_MLP_J2, true><<<g, 256, 0, s>>>(a);
    else material_kernel<FECX_MLP_J2, false><<<g, 256, 0, s>>>(a);
  } else {
    if (surrogate) material_kernel<FECX_MLP_NH, true><<<g, 256, 0, s>>>(a);
    else material_kernel<FECX_MLP_NH, false><<<g, 256, 0, s>>>(a);
  }
}
void mlp_launch_j2_paths(int64_t npaths, int steps, float max_norm, uint64_t seed, const MatParams& mp, float* X,
                         float* T, hipStream_t s) {
  if (npaths > 0) j2_paths_kernel<<<(unsigned)((npaths + 127) / 128), 128, 0, s>>>(npaths, steps, max_norm, seed, mp, X, T);
}
void mlp_launch_column_stats(const float* X, int64_t n, int w, int ncol, float* shift, float* scale, hipStream_t s) {
  column_stats_kernel<<<(unsigned)ncol, 256, 0, s>>>(X, n, w, ncol, shift, scale);
}
void mlp_launch_j2_features(const float* in, int64_t n, float* x, hipStream_t s) {
  if (n > 0) j2_feature_rows_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, n, x);
}

template <int ARCH>
static ArchInfo make_info() {
  using A = Arch<ARCH>;
  using O = Offs<A>;
  ArchInfo r{};
  r.nl = A::NL;
  for (int l = 0; l <= A::NL; ++l) { r.d[l] = A::D[l]; r.p[l] = A::P[l]; }
  r.in_w = Feat<ARCH>::IN_W;
  r.nfeat = Feat<ARCH>::NF;
  r.out_w = Feat<ARCH>::OUT_W;
  r.nparams = O::params();
  r.wf_vecs = O::wf_total();
  r.wb_vecs = O::wb_total();
  r.nbias = O::bias_total();
  r.wgrad_tiles = wgrad_tiles<ARCH>();
  return r;
}

ArchInfo arch_info(int arch) { return arch == FECX_MLP_J2 ? make_info<FECX_MLP_J2>() : make_info<FECX_MLP_NH>(); }

}  // namespace fecx
