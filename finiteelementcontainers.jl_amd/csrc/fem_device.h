// Device-side finite-element building blocks shared by the element kernels (element.hip) and the batched
// quadrature-point material kernels (material.hip): element traits, quadrature, shape functions and the FP32
// material models with their compact tangents.
#pragma once

#include <type_traits>

#include "fecx_internal.h"

namespace fecx {

template <int KIND>
struct ET;
template <>
struct ET<FECX_QUAD4> {
  static constexpr int NNPE = 4, ND = 2, NQ = 4, G = 4, NPAIRS = 10, MAXP = 3;
};
template <>
struct ET<FECX_HEX8> {
  static constexpr int NNPE = 8, ND = 3, NQ = 8, G = 8, NPAIRS = 36, MAXP = 5;
};
template <>
struct ET<FECX_TET10> {
  static constexpr int NNPE = 10, ND = 3, NQ = 4, G = 16, NPAIRS = 55, MAXP = 6;
};

__host__ __device__ constexpr int lane_cnt(int nnpe, int a) {
  return a >= nnpe ? 0 : (nnpe - 1) / 2 + 1 + ((nnpe % 2 == 0 && a < nnpe / 2) ? 1 : 0);
}

// ------------------------------------------------------------------ reference elements (FP32; FP64 forms for tet10)
template <int KIND, class R = float>
__device__ __forceinline__ R qweight() {
  if constexpr (KIND == FECX_TET10) return R(1) / R(24);
  else return R(1);
}

// Quadrature point q in reference coordinates: Gauss-Legendre 2 points per direction for quad4/hex8 (xi fastest),
// the 4-point degree-2 rule (a=(5-sqrt5)/20, b=(5+3sqrt5)/20, weights 1/24) for tet10.
template <int KIND>
__device__ __forceinline__ void qpoint(int q, float* xi) {
  constexpr float g = 0.57735026918962576f;
  if constexpr (KIND == FECX_QUAD4) {
    xi[0] = (q & 1) ? g : -g; xi[1] = (q & 2) ? g : -g;
  } else if constexpr (KIND == FECX_HEX8) {
    xi[0] = (q & 1) ? g : -g; xi[1] = (q & 2) ? g : -g; xi[2] = (q & 4) ? g : -g;
  } else {
    constexpr float qa = 0.13819660112501051f, qb = 0.58541019662496845f;
    xi[0] = q == 1 ? qb : qa; xi[1] = q == 2 ? qb : qa; xi[2] = q == 3 ? qb : qa;
  }
}
// FP64 form (the tet10 element geometry is evaluated in FP64)
template <int KIND>
__device__ __forceinline__ void qpoint(int q, double* xi) {
  static_assert(KIND == FECX_TET10, "FP64 quadrature points for tet10 only");
  constexpr double qa = 0.13819660112501051, qb = 0.58541019662496845;
  xi[0] = q == 1 ? qb : qa; xi[1] = q == 2 ? qb : qa; xi[2] = q == 3 ? qb : qa;
}

// tet10 N_a and dN_a/dxi in scalar type R
template <class R>
__device__ __forceinline__ void shape_tet10(int a, const R* xi, R& N, R* dN) {
  // barycentric L0 = 1 - x - y - z, L1 = x, L2 = y, L3 = z
  auto L = [&](int v) { return v == 0 ? R(1) - xi[0] - xi[1] - xi[2] : xi[v - 1]; };
  auto dL = [&](int v, int j) { return v == 0 ? R(-1) : (v - 1 == j ? R(1) : R(0)); };
  if (a < 4) {
    const R l = L(a);
    N = l * (2 * l - 1);
#pragma unroll
    for (int j = 0; j < 3; ++j) dN[j] = (4 * l - 1) * dL(a, j);
  } else {
    // edges 4:(0,1) 5:(1,2) 6:(2,0) 7:(0,3) 8:(1,3) 9:(2,3)
    const int m = a - 4;
    const int p = m < 3 ? m : m - 3;
    const int r = m < 3 ? (m + 1) % 3 : 3;
    const R lp = L(p), lr = L(r);
    N = 4 * lp * lr;
#pragma unroll
    for (int j = 0; j < 3; ++j) dN[j] = 4 * (dL(p, j) * lr + lp * dL(r, j));
  }
}

// N_a and dN_a/dxi of one local node a (quad4/hex8 in the structured mesh's local order, tet10 in Exodus TETRA10
// order: vertices, then edges 12, 23, 31, 14, 24, 34).
template <int KIND>
__device__ __forceinline__ void shape_node(int a, const float* xi, float& N, float* dN) {
  if constexpr (KIND == FECX_QUAD4) {
    const float sx = ((a & 3) == 1 || (a & 3) == 2) ? 1.0f : -1.0f, sy = (a & 2) ? 1.0f : -1.0f;
    const float fx = 1 + sx * xi[0], fy = 1 + sy * xi[1];
    N = 0.25f * fx * fy;
    dN[0] = 0.25f * sx * fy;
    dN[1] = 0.25f * fx * sy;
  } else if constexpr (KIND == FECX_HEX8) {
    const float sx = ((a & 3) == 1 || (a & 3) == 2) ? 1.0f : -1.0f, sy = (a & 2) ? 1.0f : -1.0f;
    const float sz = (a & 4) ? 1.0f : -1.0f;
    const float fx = 1 + sx * xi[0], fy = 1 + sy * xi[1], fz = 1 + sz * xi[2];
    N = 0.125f * fx * fy * fz;
    dN[0] = 0.125f * sx * fy * fz;
    dN[1] = 0.125f * fx * sy * fz;
    dN[2] = 0.125f * fx * fy * sz;
  } else {
    shape_tet10<float>(a, xi, N, dN);
  }
}

template <int KIND>
__device__ __forceinline__ void shape_node(int a, const double* xi, double& N, double* dN) {
  static_assert(KIND == FECX_TET10, "FP64 shape functions for tet10 only");
  shape_tet10<double>(a, xi, N, dN);
}

// log(1 + x) in FP32 without the library's ~120-instruction log1pf:
// - |x| <= 1/4: log1p(x) = 2 atanh(s) with s = x / (2 + x) (|s| <= 1/7), by the odd series
//   2 s (1 + s^2/3 + s^4/5 + s^6/7 + s^8/9), whose next term is below 4e-10 relative. 2 + x is carried as an exact
//   two-float sum and s as a two-float quotient (one residual step), so the result is within about 0.53 ulp (checked
//   against FP64 log1p on 2M points of [-1/4, 1/4] with the same operations in emulation);
// - elsewhere: Goldberg's method ("What every computer scientist should know about floating-point arithmetic", ACM
//   Computing Surveys 23, 1991, Theorem 4), u = fl(1 + x), log1p(x) = log(u) x / (u - 1), with the hardware base-2
//   logarithm and reciprocal (a few ulp).
__device__ __forceinline__ float log1p_fast(float x) {
  if (fabsf(x) <= 0.25f) {
    const float dh = 2.0f + x;
    const float dl = (2.0f - dh) + x;  // 2 + x = dh + dl exactly
    const float r = __builtin_amdgcn_rcpf(dh);
    const float sh = x * r;
    const float sl = (fmaf(-sh, dh, x) - sh * dl) * r;  // s = sh + sl: sl = (x - sh (dh + dl)) / dh
    const float z = sh * sh;
    const float q = z * fmaf(z, fmaf(z, fmaf(z, 1.0f / 9.0f, 1.0f / 7.0f), 1.0f / 5.0f), 1.0f / 3.0f);
    const float s2 = 2.0f * sh;
    return fmaf(s2, q, 2.0f * sl) + s2;
  }
  const float u = 1.0f + x;
  return __builtin_amdgcn_logf(u) * 0.693147180559945309f * (x * __builtin_amdgcn_rcpf(u - 1.0f));
}

// ------------------------------------------------------------------ materials (FP32)
// Each material provides:
//   NC                         floats of compact tangent data per quadrature point
//   eval(mp, H, so, sn, P, C)  stress P (NF x ND) and tangent data C from the field gradient H (and state); H and P
//                              in FP32 or (the tet10 element chain) FP64, C always FP32
//   prep(mp, C, g, v)          per-node vector v (NV floats, v[0..ND) = g) from the shape gradient g: K_ab depends
//                              on the pair only through v_a, v_b
//   NA, side(mp, C, w, va, s)  the a-side factors s (NA floats) of one quadrature point, weight folded in
//   pair(s, vb, K)             K += w * (g_a . A . g_b) as an NF x NF block
//   dP(mp, C, Gv, out)         out = A : Gv (matrix-free action)
//   FROM_QP                    true when P and C come precomputed per quadrature point (surrogate materials)
template <int PHYS>
struct Mat;

// -div(grad u) = f (test/poisson/TestPoissonCommon.jl:75-127): P = grad u, A = I
template <>
struct Mat<FECX_POISSON> {
  static constexpr bool FROM_QP = false;  // evaluated in the element kernel
  static constexpr int NF = 1, NC = 1, NV = 3, NS = 0, S0 = 0;  // side scalars: C[S0 .. S0+NS)
  template <int ND, class R>
  __device__ static void eval(const MatParams&, const R* H, const float*, float*, R* P, float*, bool&) {
#pragma unroll
    for (int j = 0; j < ND; ++j) P[j] = H[j];
  }
  template <int ND>
  __device__ static void prep(const MatParams&, const float*, const float* g, float* v) {
#pragma unroll
    for (int j = 0; j < ND; ++j) v[j] = g[j];
  }
  static constexpr int NA = 3;
  template <int ND>
  __device__ static void side(const MatParams&, const float*, float w, const float* va, float* sa) {
#pragma unroll
    for (int j = 0; j < ND; ++j) sa[j] = w * va[j];
  }
  template <int ND>
  __device__ static void pair(const float* sa, const float* vb, float* K) {
#pragma unroll
    for (int j = 0; j < ND; ++j) K[0] = fmaf(sa[j], vb[j], K[0]);
  }
  static constexpr int KX = 0;  // no extra accumulator word
  template <int ND>
  __device__ static void finish(float*) {}
  static constexpr bool DIAG = false;  // no cheaper form of the block K_aa
  template <int ND>
  __device__ static void dP(const MatParams&, const float*, const float* Gv, float* out) {
#pragma unroll
    for (int j = 0; j < ND; ++j) out[j] = Gv[j];
  }
};

// Small-strain linear elasticity of the reference's test Mechanics physics, psi = K/2 tr(eps)^2 + G dev eps : dev eps
// (test/mechanics/TestMechanicsCommon.jl:14-20): A = (K - 2G/3) I(x)I + G (d_ik d_jl + d_il d_jk).
template <>
struct Mat<FECX_LINEAR_ELASTIC> {
  // ND = 3, or ND = 2 in plane strain: the 2 x 2 gradient embedded in a 3 x 3 one with a zero third row and column
  // (modify_field_gradients, src/Formulations.jl:421-427), whose stress and tangent restricted to the in-plane
  // components (extract_stress / extract_stiffness, Formulations.jl:391-419) are the same formulas over ND indices
  static constexpr bool FROM_QP = false;  // evaluated in the element kernel
  static constexpr int NF = 3, NC = 1, NV = 3, NS = 0, S0 = 0;
  template <int ND, class R>
  __device__ static void eval(const MatParams& mp, const R* H, const float*, float*, R* P, float*, bool&) {
    const R K = mp.p[1], G = mp.p[2], lm = K - R(2) * G / R(3);
    R tr = R(0);
#pragma unroll
    for (int i = 0; i < ND; ++i) tr += H[i * ND + i];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int j = 0; j < ND; ++j) P[i * ND + j] = G * (H[i * ND + j] + H[j * ND + i]) + (i == j ? lm * tr : R(0));
  }
  template <int ND>
  __device__ static void prep(const MatParams&, const float*, const float* g, float* v) {
#pragma unroll
    for (int j = 0; j < ND; ++j) v[j] = g[j];
  }
  static constexpr int NA = 6;  // {w lm g_a, w G g_a}: sa[j], sa[ND + j]
  template <int ND>
  __device__ static void side(const MatParams& mp, const float*, float w, const float* ga, float* sa) {
    const float G = mp.p[2], lm = mp.p[1] - 2.0f * G / 3.0f;
#pragma unroll
    for (int j = 0; j < ND; ++j) { sa[j] = w * lm * ga[j]; sa[ND + j] = w * G * ga[j]; }
  }
  // Kb[ND * ND] accumulates the scalar sum_q w G g_a . g_b, added to the diagonal once by finish()
  template <int ND>
  __device__ static void pair(const float* sa, const float* gb, float* Kb) {
    float gg = Kb[ND * ND];
#pragma unroll
    for (int j = ND - 1; j >= 0; --j) gg = fmaf(sa[ND + j], gb[j], gg);
    Kb[ND * ND] = gg;
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int k = 0; k < ND; ++k) {
        const float t = fmaf(sa[i], gb[k], Kb[i * ND + k]);
        Kb[i * ND + k] = fmaf(gb[i], sa[ND + k], t);
      }
  }
  static constexpr int KX = 1;
  template <int ND>
  __device__ static void finish(float* Kb) {
#pragma unroll
    for (int i = 0; i < ND; ++i) Kb[i * ND + i] += Kb[ND * ND];
  }
  // the block K_aa (gb = ga): the two rank-one terms are one, g_a (w (K - 2G/3) + w G) g_a^T, and symmetric: its upper
  // triangle only (the caller mirrors it once the points are summed)
  static constexpr bool DIAG = true;
  template <int ND>
  __device__ static void pair_diag(const float* sa, const float* ga, float* Kb) {
    float gg = Kb[ND * ND];
#pragma unroll
    for (int j = ND - 1; j >= 0; --j) gg = fmaf(sa[ND + j], ga[j], gg);
    Kb[ND * ND] = gg;
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int k = i; k < ND; ++k) Kb[i * ND + k] = fmaf(ga[i], sa[k] + sa[ND + k], Kb[i * ND + k]);
  }
  template <int ND>
  __device__ static void dP(const MatParams& mp, const float*, const float* Gv, float* out) {
    const float G = mp.p[2], lm = mp.p[1] - 2.0f * G / 3.0f;
    float tr = 0.0f;
#pragma unroll
    for (int i = 0; i < ND; ++i) tr += Gv[i * ND + i];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int j = 0; j < ND; ++j) out[i * ND + j] = G * (Gv[i * ND + j] + Gv[j * ND + i]) + (i == j ? lm * tr : 0.0f);
  }
};

// Compressible neo-Hookean, psi = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2, written in the displacement gradient H
// so that no term cancels for small H: J - 1 = tr H + (tr^2 H - tr H^2)/2 + det H, ln J = log1p(J - 1),
// P = mu H + mu H^T F^-T + lambda ln J F^-T.
// A_iJkL = mu d_ik d_JL + (mu - lambda ln J) Finv_Li Finv_Jk + lambda Finv_Ji Finv_Lk, so with h = F^-1 g
// (h_i = sum_J F^-T_iJ g_J): K_ab = mu (g_a . g_b) I + (mu - lambda ln J) h_b (x) h_a + lambda h_a (x) h_b.
// C = {F^-T (9, row-major), mu - lambda ln J}.
template <>
struct Mat<FECX_NEOHOOKEAN> {
  static constexpr bool FROM_QP = false;  // evaluated in the element kernel
  static constexpr int NF = 3, NC = 10, NV = 6, NS = 1, S0 = 9;
  template <int ND, class R>
  __device__ static void eval(const MatParams& mp, const R* H, const float*, float*, R* P, float* C, bool& bad) {
    const R lam = mp.p[0], mu = mp.p[1];
    const R tr = H[0] + H[4] + H[8];
    const R trH2 = H[0] * H[0] + H[4] * H[4] + H[8] * H[8] + R(2) * (H[1] * H[3] + H[2] * H[6] + H[5] * H[7]);
    const R detH = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                   H[2] * (H[3] * H[7] - H[4] * H[6]);
    const R Jm1 = tr + R(0.5) * (tr * tr - trH2) + detH;
    const R J = R(1) + Jm1;
    if (!(J > R(0))) bad = true;
    R lnJ;
    if constexpr (std::is_same_v<R, float>) lnJ = log1p_fast(Jm1);
    else lnJ = log1p(Jm1);
    const R F[9] = {1 + H[0], H[1], H[2], H[3], 1 + H[4], H[5], H[6], H[7], 1 + H[8]};
    const R iJ = R(1) / J;
    R FiT[9];  // cofactor(F) / J
    FiT[0] = (F[4] * F[8] - F[5] * F[7]) * iJ; FiT[1] = (F[5] * F[6] - F[3] * F[8]) * iJ;
    FiT[2] = (F[3] * F[7] - F[4] * F[6]) * iJ; FiT[3] = (F[2] * F[7] - F[1] * F[8]) * iJ;
    FiT[4] = (F[0] * F[8] - F[2] * F[6]) * iJ; FiT[5] = (F[1] * F[6] - F[0] * F[7]) * iJ;
    FiT[6] = (F[1] * F[5] - F[2] * F[4]) * iJ; FiT[7] = (F[2] * F[3] - F[0] * F[5]) * iJ;
    FiT[8] = (F[0] * F[4] - F[1] * F[3]) * iJ;
    if (P) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const R htf = H[0 * 3 + i] * FiT[0 * 3 + j] + H[1 * 3 + i] * FiT[1 * 3 + j] + H[2 * 3 + i] * FiT[2 * 3 + j];
          P[i * 3 + j] = mu * (H[i * 3 + j] + htf) + lam * lnJ * FiT[i * 3 + j];
        }
    }
    if (C) {
#pragma unroll
      for (int i = 0; i < 9; ++i) C[i] = (float)FiT[i];
      C[9] = (float)(mu - lam * lnJ);
    }
  }
  template <int ND>
  __device__ static void prep(const MatParams&, const float* C, const float* g, float* v) {
    v[0] = g[0]; v[1] = g[1]; v[2] = g[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[3 + i] = C[i * 3 + 0] * g[0] + C[i * 3 + 1] * g[1] + C[i * 3 + 2] * g[2];
  }
  static constexpr int NA = 9;  // {w mu g_a, w c1 h_a, w lambda h_a}
  template <int ND>
  __device__ static void side(const MatParams& mp, const float* C, float w, const float* va, float* sa) {
    const float wl = w * mp.p[0], wm = w * mp.p[1], wc = w * C[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) { sa[j] = wm * va[j]; sa[3 + j] = wc * va[3 + j]; sa[6 + j] = wl * va[3 + j]; }
  }
  // Kb[9] accumulates the scalar sum_q w mu g_a . g_b, added to the diagonal once by finish()
  template <int ND>
  __device__ static void pair(const float* sa, const float* vb, float* Kb) {
    Kb[9] = fmaf(sa[0], vb[0], fmaf(sa[1], vb[1], fmaf(sa[2], vb[2], Kb[9])));
    const float* hb = vb + 3;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float t = fmaf(hb[i], sa[3 + k], Kb[i * 3 + k]);
        Kb[i * 3 + k] = fmaf(sa[6 + i], hb[k], t);
      }
  }
  static constexpr int KX = 1;
  template <int ND>
  __device__ static void finish(float* Kb) {
#pragma unroll
    for (int i = 0; i < 3; ++i) Kb[i * 3 + i] += Kb[9];
  }
  // the block K_aa (vb = va): the two rank-one terms are one, (c1 + lambda) w h_a h_a^T, and symmetric: its upper
  // triangle only (the caller mirrors it once the points are summed)
  static constexpr bool DIAG = true;
  template <int ND>
  __device__ static void pair_diag(const float* sa, const float* va, float* Kb) {
    Kb[9] = fmaf(sa[0], va[0], fmaf(sa[1], va[1], fmaf(sa[2], va[2], Kb[9])));
    const float* ha = va + 3;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = i; k < 3; ++k) Kb[i * 3 + k] = fmaf(ha[i], sa[3 + k] + sa[6 + k], Kb[i * 3 + k]);
  }
  template <int ND>
  __device__ static void dP(const MatParams& mp, const float* C, const float* Gv, float* out) {
    // mu Gv + c1 F^-T Gv^T F^-T + lambda (F^-T : Gv) F^-T
    const float lam = mp.p[0], mu = mp.p[1], c1 = C[9];
    float M[9];  // Gv^T F^-T
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int j = 0; j < 3; ++j) M[l * 3 + j] = Gv[0 * 3 + l] * C[0 * 3 + j] + Gv[1 * 3 + l] * C[1 * 3 + j] + Gv[2 * 3 + l] * C[2 * 3 + j];
    float tr = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) tr += C[i] * Gv[i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        out[i * 3 + j] = mu * Gv[i * 3 + j] + c1 * (C[i * 3 + 0] * M[0 * 3 + j] + C[i * 3 + 1] * M[1 * 3 + j] + C[i * 3 + 2] * M[2 * 3 + j]) +
                         lam * tr * C[i * 3 + j];
  }
};

// J2 small-strain plasticity with linear isotropic hardening, one-step radial return and its consistent tangent
// (Simo & Hughes, Computational Inelasticity, Box 3.1-3.2). State (7): plastic strain in SymmetricTensor order
// (11,21,31,22,32,33), then equivalent plastic strain. Tangent C = kappa I(x)I + 2 mu theta (I_sym - I(x)I/3)
// - 2 mu thetabar n(x)n, so with m = n g: K_ab = b1 g_a (x) g_b + b2 ((g_a . g_b) I + g_b (x) g_a) - b3 m_a (x) m_b,
// b1 = kappa - 2 mu theta / 3, b2 = mu theta, b3 = 2 mu thetabar. C = {n (11,22,33,12,23,13), b1, b2, b3}.
// p = {E, nu, sigma_y, H, mu, kappa, 1 / (2 mu + 2 H / 3), 1 / (1 + H / (3 mu))}: the last four formed on the host
// (fecx_set_material), so that the point update divides once (by |s|) instead of ten times.
__device__ __forceinline__ void j2_update(const float* p, const float* H, const float* so, float* sn, float* P, float* C) {
  const float sy = p[2], Hh = p[3], mu = p[4], kap = p[5], idg = p[6], ithb = p[7];
  constexpr float third = 1.0f / 3.0f;
  const float ep[9] = {so[0], so[1], so[2], so[1], so[3], so[4], so[2], so[4], so[5]};
  float ee[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) ee[i * 3 + j] = 0.5f * (H[i * 3 + j] + H[j * 3 + i]) - ep[i * 3 + j];
  const float tr = ee[0] + ee[4] + ee[8];
  float s[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) s[i] = 2 * mu * ee[i];
  const float sv = 2 * mu * (tr * third);
  s[0] -= sv; s[4] -= sv; s[8] -= sv;
  float sn2 = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) sn2 += s[i] * s[i];
  const float snorm = sqrtf(sn2);
  const float r23 = 0.81649658092772603f;
  const float f = snorm - r23 * (sy + Hh * so[6]);
  const bool plastic = f > 0.0f;
  const float inv = plastic ? 1.0f / snorm : 0.0f;
  float nh[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) nh[i] = s[i] * inv;
  const float dg = plastic ? f * idg : 0.0f;
  if (P) {
#pragma unroll
    for (int i = 0; i < 9; ++i) P[i] = s[i] - 2 * mu * dg * nh[i];
    P[0] += kap * tr; P[4] += kap * tr; P[8] += kap * tr;
  }
  if (sn) {
    sn[0] = ep[0] + dg * nh[0]; sn[1] = ep[3] + dg * nh[3]; sn[2] = ep[6] + dg * nh[6];
    sn[3] = ep[4] + dg * nh[4]; sn[4] = ep[7] + dg * nh[7]; sn[5] = ep[8] + dg * nh[8];
    sn[6] = so[6] + r23 * dg;
  }
  if (C) {
    const float th = plastic ? 1.0f - 2 * mu * dg * inv : 1.0f;
    const float thb = plastic ? ithb - (1.0f - th) : 0.0f;
    C[0] = nh[0]; C[1] = nh[4]; C[2] = nh[8]; C[3] = nh[1]; C[4] = nh[5]; C[5] = nh[2];
    C[6] = kap - 2 * mu * th * third; C[7] = mu * th; C[8] = 2 * mu * thb;
  }
}

template <>
struct Mat<FECX_J2> {
  static constexpr bool FROM_QP = false;  // evaluated in the element kernel
  static constexpr int NF = 3, NC = 9, NV = 6, NS = 3, S0 = 6;
  template <int ND, class R>
  __device__ static void eval(const MatParams& mp, const R* H, const float* so, float* sn, R* P, float* C, bool&) {
    if constexpr (std::is_same_v<R, float>) {
      j2_update(mp.p, H, so, sn, P, C);
    } else {  // the FP64 element chain: the return mapping itself stays FP32
      float Hf[9], Pf[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) Hf[c] = (float)H[c];
      j2_update(mp.p, Hf, so, sn, Pf, C);
#pragma unroll
      for (int c = 0; c < 9; ++c) P[c] = Pf[c];
    }
  }
  template <int ND>
  __device__ static void prep(const MatParams&, const float* C, const float* g, float* v) {
    v[0] = g[0]; v[1] = g[1]; v[2] = g[2];
    v[3] = C[0] * g[0] + C[3] * g[1] + C[5] * g[2];
    v[4] = C[3] * g[0] + C[1] * g[1] + C[4] * g[2];
    v[5] = C[5] * g[0] + C[4] * g[1] + C[2] * g[2];
  }
  static constexpr int NA = 9;  // {w b1 g_a, w b2 g_a, -w b3 m_a}
  template <int ND>
  __device__ static void side(const MatParams&, const float* C, float w, const float* va, float* sa) {
    const float w1 = w * C[6], w2 = w * C[7], w3 = -w * C[8];
#pragma unroll
    for (int j = 0; j < 3; ++j) { sa[j] = w1 * va[j]; sa[3 + j] = w2 * va[j]; sa[6 + j] = w3 * va[3 + j]; }
  }
  // the isotropic term added to the diagonal at every point: a separate accumulator (as the other vector materials
  // keep) measured 0.369 / 0.372 -> 0.376 / 0.374 ms on the fused J2 hex8 64^3 kernel (more spilled registers)
  template <int ND>
  __device__ static void pair(const float* sa, const float* vb, float* Kb) {
    const float gg = fmaf(sa[3], vb[0], fmaf(sa[4], vb[1], sa[5] * vb[2]));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float t = fmaf(sa[i], vb[k], Kb[i * 3 + k]);
        t = fmaf(vb[i], sa[3 + k], t);
        t = fmaf(sa[6 + i], vb[3 + k], t);
        Kb[i * 3 + k] = i == k ? t + gg : t;
      }
  }
  static constexpr int KX = 0;
  template <int ND>
  __device__ static void finish(float*) {}
  // the block K_aa (vb = va): g_a (w b1 + w b2) g_a^T - w b3 m_a m_a^T + the isotropic term, symmetric: its upper
  // triangle only (the caller mirrors it once the points are summed)
  static constexpr bool DIAG = true;
  template <int ND>
  __device__ static void pair_diag(const float* sa, const float* va, float* Kb) {
    const float gg = fmaf(sa[3], va[0], fmaf(sa[4], va[1], sa[5] * va[2]));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = i; k < 3; ++k) {
        float t = fmaf(va[i], sa[k] + sa[3 + k], Kb[i * 3 + k]);
        t = fmaf(sa[6 + i], va[3 + k], t);
        Kb[i * 3 + k] = i == k ? t + gg : t;
      }
  }
  template <int ND>
  __device__ static void dP(const MatParams&, const float* C, const float* Gv, float* out) {
    const float b1 = C[6], b2 = C[7], b3 = C[8];
    const float tr = Gv[0] + Gv[4] + Gv[8];
    const float n[9] = {C[0], C[3], C[5], C[3], C[1], C[4], C[5], C[4], C[2]};
    float nd = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) nd += n[i] * Gv[i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        out[i * 3 + j] = b2 * (Gv[i * 3 + j] + Gv[j * 3 + i]) + (i == j ? b1 * tr : 0.0f) - b3 * nd * n[i * 3 + j];
  }
};

// Surrogate materials: P and the compact tangent (same format as the plain model's) are computed per quadrature point
// by the batched material pass (mlp.hip: network, residual check, plain fallback) and read by the element kernel.
template <>
struct Mat<FECX_J2_SURROGATE> : Mat<FECX_J2> {
  static constexpr bool FROM_QP = true;
};
template <>
struct Mat<FECX_NH_SURROGATE> : Mat<FECX_NEOHOOKEAN> {
  static constexpr bool FROM_QP = true;
};

// ------------------------------------------------------------------ LDS staging shared by the element and mass kernels
typedef float f32x4 __attribute__((ext_vector_type(4)));  // 16-byte moves (nontemporal stores need a vector type)

// One element's quadrature points, padded to a stride of 8 (mod 16) words: the 8 elements of a hex8 wavefront (4 of a
// tet10 one) then start on distinct LDS banks when their node lanes read the same point (J2 hex8, 352 words unpadded,
// put all 8 on two bank offsets: its element kernel ran 7 % slower than with the per-element layout).
template <class QP, int NQ, int PADW>
struct QPBlock {
  QP q[NQ];
  float pad[PADW];
};
template <class QP, int NQ>
struct QPBlock<QP, NQ, 0> {
  QP q[NQ];
};
template <class QP, int NQ>
constexpr int qp_pad() {
  constexpr int w = (int)(NQ * sizeof(QP) / sizeof(float));
  return ((8 - w % 16) % 16 + 16) % 16;
}
}  // namespace fecx
