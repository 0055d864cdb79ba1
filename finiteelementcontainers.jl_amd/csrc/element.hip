// Element kernels (phase 1 of assembly): FP32 element math, element-private FP32 outputs.
//
// Reference: the generic element loop _assemble_block! (src/assemblers/Assemblers.jl:402-432) runs one
// work-item per element and stores K_el into element-private COO slots (Assemblers.jl:109-124). Here one
// wavefront processes 64/G elements with G lanes per element:
//   gather   lane a  (a < nnpe) : connectivity, coordinates and nodal values of local node a; element-local
//                                 differences x_a - x_0 (hi/lo FP32 coordinate split) and u_a - u_0
//   qp       lane q  (q < nq)   : isoparametric map, shape gradients, field gradient and material point
//                                 (stress, symmetric tangent) of quadrature point q, staged in LDS
//   node     lane a  (a < nnpe) : residual / action rows of node a, and the node-pair blocks K_ab it owns
//                                 (symmetric storage, one NF x NF block per unordered pair)
// Outputs are FP32 element-private slots; the FP64 sums into the residual field and the CSR values happen
// in the reduction kernels (reduce.hip), through the precomputed inverse COO->CSR map.
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
__host__ __device__ constexpr int lane_off(int nnpe, int a) {
  return a * ((nnpe - 1) / 2 + 1) + ((nnpe % 2 == 0) ? (a < nnpe / 2 ? a : nnpe / 2) : 0);
}

// ------------------------------------------------------------------ reference elements (FP32)
template <int KIND>
__device__ __forceinline__ float qweight(int q) {
  if constexpr (KIND == FECX_TET10) return 1.0f / 24.0f;
  else return 1.0f;
}

// N[a], dN[a][j] at quadrature point q. Gauss-Legendre 2 points per direction for quad4/hex8 (xi fastest),
// the 4-point degree-2 rule (a=(5-sqrt5)/20, b=(5+3sqrt5)/20) for tet10.
template <int KIND>
__device__ __forceinline__ void shape_at(int q, float* N, float (*dN)[ET<KIND>::ND]) {
  constexpr float g = 0.57735026918962576f;
  if constexpr (KIND == FECX_QUAD4) {
    const float xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g;
    const float s[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float fx = 1 + s[a][0] * xi, fy = 1 + s[a][1] * eta;
      N[a] = 0.25f * fx * fy;
      dN[a][0] = 0.25f * s[a][0] * fy;
      dN[a][1] = 0.25f * fx * s[a][1];
    }
  } else if constexpr (KIND == FECX_HEX8) {
    const float xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g, zeta = (q & 4) ? g : -g;
    const float s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      float fx = 1 + s[a][0] * xi, fy = 1 + s[a][1] * eta, fz = 1 + s[a][2] * zeta;
      N[a] = 0.125f * fx * fy * fz;
      dN[a][0] = 0.125f * s[a][0] * fy * fz;
      dN[a][1] = 0.125f * fx * s[a][1] * fz;
      dN[a][2] = 0.125f * fx * fy * s[a][2];
    }
  } else {
    constexpr float qa = 0.13819660112501051f, qb = 0.58541019662496845f;
    float p[3] = {qa, qa, qa};
    if (q >= 1) p[q - 1] = qb;
    const float L[4] = {1.0f - p[0] - p[1] - p[2], p[0], p[1], p[2]};
    const float dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      N[a] = L[a] * (2 * L[a] - 1);
#pragma unroll
      for (int j = 0; j < 3; ++j) dN[a][j] = (4 * L[a] - 1) * dL[a][j];
    }
    const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int P = ed[m][0], Q = ed[m][1];
      N[4 + m] = 4 * L[P] * L[Q];
#pragma unroll
      for (int j = 0; j < 3; ++j) dN[4 + m][j] = 4 * (dL[P][j] * L[Q] + L[P] * dL[Q][j]);
    }
  }
}

// ------------------------------------------------------------------ materials (FP32)
__device__ __forceinline__ int asym_idx(int r, int c) {
  // upper triangle of a 9 x 9 symmetric matrix, row-major
  return r <= c ? 8 * r - (r * (r - 1)) / 2 + c : 8 * c - (c * (c - 1)) / 2 + r;
}

// Neo-Hookean, psi = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2, written in the displacement gradient H so
// that no term cancels for small H: J - 1 = tr H + (tr^2 H - tr H^2)/2 + det H, ln J = log1p(J - 1),
// P = mu H + mu H^T F^-T + lambda ln J F^-T.
__device__ __forceinline__ void neohookean(float lam, float mu, const float* H, float* P, float* A, bool& bad) {
  const float tr = H[0] + H[4] + H[8];
  const float trH2 = H[0] * H[0] + H[4] * H[4] + H[8] * H[8] + 2.0f * (H[1] * H[3] + H[2] * H[6] + H[5] * H[7]);
  const float detH = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                     H[2] * (H[3] * H[7] - H[4] * H[6]);
  const float Jm1 = tr + 0.5f * (tr * tr - trH2) + detH;
  const float J = 1.0f + Jm1;
  if (!(J > 0.0f)) bad = true;
  const float lnJ = log1pf(Jm1);
  const float F[9] = {1 + H[0], H[1], H[2], H[3], 1 + H[4], H[5], H[6], H[7], 1 + H[8]};
  // cofactor(F) = J F^-T
  float C[9];
  C[0] = F[4] * F[8] - F[5] * F[7]; C[1] = F[5] * F[6] - F[3] * F[8]; C[2] = F[3] * F[7] - F[4] * F[6];
  C[3] = F[2] * F[7] - F[1] * F[8]; C[4] = F[0] * F[8] - F[2] * F[6]; C[5] = F[1] * F[6] - F[0] * F[7];
  C[6] = F[1] * F[5] - F[2] * F[4]; C[7] = F[2] * F[3] - F[0] * F[5]; C[8] = F[0] * F[4] - F[1] * F[3];
  const float iJ = 1.0f / J;
  float FiT[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) FiT[i] = C[i] * iJ;
  if (P) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        float htf = H[0 * 3 + i] * FiT[0 * 3 + j] + H[1 * 3 + i] * FiT[1 * 3 + j] + H[2 * 3 + i] * FiT[2 * 3 + j];
        P[i * 3 + j] = mu * (H[i * 3 + j] + htf) + lam * lnJ * FiT[i * 3 + j];
      }
  }
  if (A) {
    const float c1 = mu - lam * lnJ;
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
      for (int c = r; c < 9; ++c) {
        const int i = r / 3, J_ = r % 3, k = c / 3, L = c % 3;
        float v = c1 * FiT[i * 3 + L] * FiT[k * 3 + J_] + lam * FiT[i * 3 + J_] * FiT[k * 3 + L];
        if (i == k && J_ == L) v += mu;
        A[asym_idx(r, c)] = v;
      }
  }
}

// Small-strain linear elasticity of the reference's test Mechanics physics: psi = K/2 tr(eps)^2 + G dev eps : dev eps
// (test/mechanics/TestMechanicsCommon.jl:14-20).
__device__ __forceinline__ void linear_elastic(float K, float G, const float* H, float* P, float* A) {
  const float tr = H[0] + H[4] + H[8];
  const float lm = K - 2.0f * G / 3.0f;
  if (P) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i * 3 + j] = G * (H[i * 3 + j] + H[j * 3 + i]) + (i == j ? lm * tr : 0.0f);
  }
  if (A) {
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
      for (int c = r; c < 9; ++c) {
        const int i = r / 3, j = r % 3, k = c / 3, l = c % 3;
        A[asym_idx(r, c)] = lm * (i == j) * (k == l) + G * ((i == k) * (j == l) + (i == l) * (j == k));
      }
  }
}

// J2 small-strain plasticity with linear isotropic hardening, one-step radial return and its consistent
// tangent (Simo & Hughes, Computational Inelasticity, Box 3.1-3.2). State (7): plastic strain in
// SymmetricTensor order (11,21,31,22,32,33), then equivalent plastic strain.
__device__ __forceinline__ void j2(const float* p, const float* H, const float* so, float* sn, float* P, float* A) {
  const float E = p[0], nu = p[1], sy = p[2], Hh = p[3];
  const float mu = E / (2 * (1 + nu)), kap = E / (3 * (1 - 2 * nu));
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
  s[0] -= 2 * mu * tr / 3; s[4] -= 2 * mu * tr / 3; s[8] -= 2 * mu * tr / 3;
  float sn2 = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) sn2 += s[i] * s[i];
  const float snorm = sqrtf(sn2);
  const float r23 = 0.81649658092772603f;
  const float f = snorm - r23 * (sy + Hh * so[6]);
  float dg = 0.0f, nh[9];
  const float inv = f > 0.0f ? 1.0f / snorm : 0.0f;
#pragma unroll
  for (int i = 0; i < 9; ++i) nh[i] = s[i] * inv;
  if (f > 0.0f) dg = f / (2 * mu + 2.0f * Hh / 3.0f);
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
  if (A) {
    float th = 1.0f, thb = 0.0f;
    if (f > 0.0f) {
      th = 1.0f - 2 * mu * dg / snorm;
      thb = 1.0f / (1.0f + Hh / (3 * mu)) - (1.0f - th);
    }
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
      for (int c = r; c < 9; ++c) {
        const int i = r / 3, j = r % 3, k = c / 3, l = c % 3;
        const float dij = i == j, dkl = k == l;
        A[asym_idx(r, c)] = kap * dij * dkl +
                            2 * mu * th * (0.5f * ((i == k) * (j == l) + (i == l) * (j == k)) - dij * dkl / 3.0f) -
                            2 * mu * thb * nh[r] * nh[c];
      }
  }
}

// ------------------------------------------------------------------ the kernel
template <int KIND, int NF, int PHYS>
struct ElemLDS {
  using T = ET<KIND>;
  static constexpr int D = NF * T::ND;
  static constexpr int AS = (PHYS == FECX_POISSON) ? 1 : 45;
  float x[T::NNPE][T::ND];
  float x0[T::ND];
  float u[T::NNPE][NF];
  float v[T::NNPE][NF];
  float kv[T::NNPE][NF];
  struct QP {
    float g[T::NNPE][T::ND];
    float w;
    float P[D];
    float src[NF];
    float A[AS];
  } qp[T::NQ];
};

template <int KIND, int NF, int PHYS, int MODE>
__global__ __launch_bounds__(64) void element_kernel(ElementArgs args) {
  using T = ET<KIND>;
  constexpr int ND = T::ND, NNPE = T::NNPE, NQ = T::NQ, G = T::G, EPW = 64 / G, NDOF = NNPE * NF;
  constexpr int D = NF * ND;
  constexpr bool DO_R = (MODE & MODE_R) != 0, DO_K = (MODE & MODE_K) != 0;
  constexpr bool DO_MF = (MODE & MODE_ACTION_MF) != 0, DO_ASM = (MODE & MODE_ACTION_ASM) != 0;
  constexpr bool NEED_A = DO_K || DO_MF || DO_ASM;
  __shared__ ElemLDS<KIND, NF, PHYS> lds[EPW];

  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const int64_t e = (int64_t)blockIdx.x * EPW + grp;
  const bool valid = e < args.ne;
  auto& L = lds[grp];
  bool bad = false;

  // ---------------- gather (lane = local node)
  {
    float hi[ND] = {}, lo[ND] = {};
    double uu[NF] = {}, vv[NF] = {};
    if (valid && lane < NNPE) {
      const int64_t node = args.conn[e * NNPE + lane];
#pragma unroll
      for (int d = 0; d < ND; ++d) { hi[d] = args.xhi[node * ND + d]; lo[d] = args.xlo[node * ND + d]; }
#pragma unroll
      for (int i = 0; i < NF; ++i) {
        uu[i] = args.U[node * NF + i];
        if (DO_MF || DO_ASM) vv[i] = args.V[node * NF + i];
      }
    }
    const int src0 = grp * G;  // lane of local node 0
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const float h0 = __shfl(hi[d], src0), l0 = __shfl(lo[d], src0);
      if (lane < NNPE) L.x[lane][d] = (hi[d] - h0) + (lo[d] - l0);
      if (lane == 0) L.x0[d] = h0 + l0;
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      const double u0 = __shfl(uu[i], src0);
      if (lane < NNPE) L.u[lane][i] = (float)(uu[i] - u0);
      if (DO_MF) {
        const double v0 = __shfl(vv[i], src0);
        if (lane < NNPE) L.v[lane][i] = (float)(vv[i] - v0);
      }
      if (DO_ASM && lane < NNPE) { L.v[lane][i] = (float)vv[i]; L.kv[lane][i] = 0.0f; }
    }
  }
  __syncthreads();

  // ---------------- quadrature points (lane = q)
  if (valid && lane < NQ) {
    const int q = lane;
    float N[NNPE], dN[NNPE][ND];
    shape_at<KIND>(q, N, dN);
    float J[ND][ND] = {};
#pragma unroll
    for (int a = 0; a < NNPE; ++a)
#pragma unroll
      for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) J[i][j] += L.x[a][i] * dN[a][j];
    float Ji[ND][ND], det;
    if constexpr (ND == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const float id = 1.0f / det;
      Ji[0][0] = J[1][1] * id; Ji[0][1] = -J[0][1] * id; Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
    } else {
      const float c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                  c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const float id = 1.0f / det;
      Ji[0][0] = c00 * id; Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[1][0] = c01 * id; Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[2][0] = c02 * id; Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
    if (!(det > 0.0f)) bad = true;
    auto& Q = L.qp[q];
    float g[NNPE][ND];
#pragma unroll
    for (int a = 0; a < NNPE; ++a)
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        float s = 0;
#pragma unroll
        for (int k = 0; k < ND; ++k) s += dN[a][k] * Ji[k][j];
        g[a][j] = s;
        Q.g[a][j] = s;
      }
    Q.w = det * qweight<KIND>(q);
    float H[D] = {};
#pragma unroll
    for (int a = 0; a < NNPE; ++a)
#pragma unroll
      for (int i = 0; i < NF; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) H[i * ND + j] += L.u[a][i] * g[a][j];
    float P[D], A[(PHYS == FECX_POISSON) ? 1 : 45];
    if constexpr (PHYS == FECX_POISSON) {
#pragma unroll
      for (int j = 0; j < ND; ++j) P[j] = H[j];
      if (DO_R) {
        float f = args.mp.p[1];
        if (args.mp.source_kind == 1) {
          float Xq[ND];
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            float s = 0;
#pragma unroll
            for (int a = 0; a < NNPE; ++a) s += N[a] * L.x[a][d];
            Xq[d] = L.x0[d] + s;
          }
          f *= (float)ND * 9.8696044010893586f;
#pragma unroll
          for (int d = 0; d < ND; ++d) f *= sinpif(Xq[d]);
        }
        Q.src[0] = f;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NF; ++i) Q.src[i] = 0.0f;
      if constexpr (PHYS == FECX_NEOHOOKEAN) {
        neohookean(args.mp.p[0], args.mp.p[1], H, P, NEED_A ? A : nullptr, bad);
      } else if constexpr (PHYS == FECX_LINEAR_ELASTIC) {
        linear_elastic(args.mp.p[1], args.mp.p[2], H, P, NEED_A ? A : nullptr);
      } else if constexpr (PHYS == FECX_J2) {
        float so[7], sn[7];
        const int64_t sidx = (e * NQ + q) * 7;
#pragma unroll
        for (int s = 0; s < 7; ++s) so[s] = args.st_old[sidx + s];
        j2(args.mp.p, H, so, sn, P, NEED_A ? A : nullptr);
        if (DO_R && args.st_new) {
#pragma unroll
          for (int s = 0; s < 7; ++s) args.st_new[sidx + s] = sn[s];
        }
      }
      if (NEED_A) {
#pragma unroll
        for (int i = 0; i < 45; ++i) Q.A[i] = A[i];
      }
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (!(P[c] == P[c])) bad = true;
    if (DO_R) {
#pragma unroll
      for (int c = 0; c < D; ++c) Q.P[c] = P[c];
    }
    if (DO_MF) {
      float Gv[D] = {};
#pragma unroll
      for (int a = 0; a < NNPE; ++a)
#pragma unroll
        for (int i = 0; i < NF; ++i)
#pragma unroll
          for (int j = 0; j < ND; ++j) Gv[i * ND + j] += L.v[a][i] * g[a][j];
#pragma unroll
      for (int r = 0; r < D; ++r) {
        float s = 0;
        if constexpr (PHYS == FECX_POISSON) s = Gv[r];
        else {
#pragma unroll
          for (int c = 0; c < D; ++c) s += A[asym_idx(r, c)] * Gv[c];
        }
        Q.P[r] = s;
      }
    }
  }
  __syncthreads();

  // ---------------- node rows (lane = a)
  if (valid && lane < NNPE) {
    const int a = lane;
    if (DO_R || DO_MF) {
      float out[NF] = {};
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const auto& Q = L.qp[q];
        float Nq[NNPE], dNq[NNPE][ND];
        if (DO_R && PHYS == FECX_POISSON) shape_at<KIND>(q, Nq, dNq);
#pragma unroll
        for (int i = 0; i < NF; ++i) {
          float s = 0;
#pragma unroll
          for (int j = 0; j < ND; ++j) s += Q.P[i * ND + j] * Q.g[a][j];
          if (DO_R && PHYS == FECX_POISSON) s -= Nq[a] * Q.src[i];
          out[i] += Q.w * s;
        }
      }
#pragma unroll
      for (int i = 0; i < NF; ++i) args.Rel[e * NDOF + a * NF + i] = out[i];
    }
    if (DO_K || DO_ASM) {
      constexpr int MAXP = T::MAXP;
      const int cnt = lane_cnt(NNPE, a);
      float acc[MAXP][NF * NF];
#pragma unroll
      for (int p = 0; p < MAXP; ++p)
#pragma unroll
        for (int c = 0; c < NF * NF; ++c) acc[p][c] = 0.0f;
#pragma unroll 1
      for (int q = 0; q < NQ; ++q) {
        const auto& Q = L.qp[q];
        float ga[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) ga[j] = Q.g[a][j];
        if constexpr (PHYS == FECX_POISSON) {
#pragma unroll
          for (int p = 0; p < MAXP; ++p) {
            if (p < cnt) {
              const int b = (a + p) % NNPE;
              float s = 0;
#pragma unroll
              for (int j = 0; j < ND; ++j) s += ga[j] * Q.g[b][j];
              acc[p][0] += Q.w * s;
            }
          }
        } else {
          // T[i][k][l] = sum_j g_a[j] A[(i,j),(k,l)]
          float Tm[NF][NF][ND];
#pragma unroll
          for (int i = 0; i < NF; ++i)
#pragma unroll
            for (int k = 0; k < NF; ++k)
#pragma unroll
              for (int l = 0; l < ND; ++l) {
                float s = 0;
#pragma unroll
                for (int j = 0; j < ND; ++j) s += ga[j] * Q.A[asym_idx(i * ND + j, k * ND + l)];
                Tm[i][k][l] = Q.w * s;
              }
#pragma unroll
          for (int p = 0; p < MAXP; ++p) {
            if (p < cnt) {
              const int b = (a + p) % NNPE;
              float gb[ND];
#pragma unroll
              for (int l = 0; l < ND; ++l) gb[l] = Q.g[b][l];
#pragma unroll
              for (int i = 0; i < NF; ++i)
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                  float s = 0;
#pragma unroll
                  for (int l = 0; l < ND; ++l) s += Tm[i][k][l] * gb[l];
                  acc[p][i * NF + k] += s;
                }
            }
          }
        }
      }
      if (DO_K) {
        float* dst = args.Kel + (e * T::NPAIRS + lane_off(NNPE, a)) * (NF * NF);
#pragma unroll
        for (int p = 0; p < MAXP; ++p)
          if (p < cnt)
#pragma unroll
            for (int c = 0; c < NF * NF; ++c) dst[p * NF * NF + c] = acc[p][c];
      }
      if (DO_ASM) {
#pragma unroll
        for (int p = 0; p < MAXP; ++p) {
          if (p < cnt) {
            const int b = (a + p) % NNPE;
#pragma unroll
            for (int i = 0; i < NF; ++i) {
              float s = 0;
#pragma unroll
              for (int k = 0; k < NF; ++k) s += acc[p][i * NF + k] * L.v[b][k];
              atomicAdd(&L.kv[a][i], s);
            }
            if (b != a) {
#pragma unroll
              for (int k = 0; k < NF; ++k) {
                float s = 0;
#pragma unroll
                for (int i = 0; i < NF; ++i) s += acc[p][i * NF + k] * L.v[a][i];
                atomicAdd(&L.kv[b][k], s);
              }
            }
          }
        }
      }
    }
  }
  if (DO_ASM) {
    __syncthreads();
    if (valid && lane < NNPE) {
#pragma unroll
      for (int i = 0; i < NF; ++i) args.Rel[e * NDOF + lane * NF + i] = L.kv[lane][i];
    }
  }
  if (bad) atomicOr(args.flag, 1);
}

template <int KIND, int NF, int PHYS>
static void launch_modes(int mode, const ElementArgs& a, hipStream_t s) {
  using T = ET<KIND>;
  constexpr int EPW = 64 / T::G;
  const unsigned grid = (unsigned)((a.ne + EPW - 1) / EPW);
  switch (mode) {
    case MODE_R: element_kernel<KIND, NF, PHYS, MODE_R><<<grid, 64, 0, s>>>(a); break;
    case MODE_K: element_kernel<KIND, NF, PHYS, MODE_K><<<grid, 64, 0, s>>>(a); break;
    case MODE_R | MODE_K: element_kernel<KIND, NF, PHYS, MODE_R | MODE_K><<<grid, 64, 0, s>>>(a); break;
    case MODE_ACTION_MF: element_kernel<KIND, NF, PHYS, MODE_ACTION_MF><<<grid, 64, 0, s>>>(a); break;
    case MODE_ACTION_ASM: element_kernel<KIND, NF, PHYS, MODE_ACTION_ASM><<<grid, 64, 0, s>>>(a); break;
    default: break;
  }
}

int launch_element(int kind, int nf, int mode, const ElementArgs& a, hipStream_t s) {
  const int phys = a.mp.phys;
  if (nf == 1 && phys == FECX_POISSON) {
    if (kind == FECX_QUAD4) { launch_modes<FECX_QUAD4, 1, FECX_POISSON>(mode, a, s); return 0; }
    if (kind == FECX_HEX8) { launch_modes<FECX_HEX8, 1, FECX_POISSON>(mode, a, s); return 0; }
    if (kind == FECX_TET10) { launch_modes<FECX_TET10, 1, FECX_POISSON>(mode, a, s); return 0; }
    return -1;
  }
  if (nf == 3 && (kind == FECX_HEX8 || kind == FECX_TET10)) {
    if (kind == FECX_HEX8) {
      if (phys == FECX_LINEAR_ELASTIC) { launch_modes<FECX_HEX8, 3, FECX_LINEAR_ELASTIC>(mode, a, s); return 0; }
      if (phys == FECX_NEOHOOKEAN) { launch_modes<FECX_HEX8, 3, FECX_NEOHOOKEAN>(mode, a, s); return 0; }
      if (phys == FECX_J2) { launch_modes<FECX_HEX8, 3, FECX_J2>(mode, a, s); return 0; }
    } else {
      if (phys == FECX_LINEAR_ELASTIC) { launch_modes<FECX_TET10, 3, FECX_LINEAR_ELASTIC>(mode, a, s); return 0; }
      if (phys == FECX_NEOHOOKEAN) { launch_modes<FECX_TET10, 3, FECX_NEOHOOKEAN>(mode, a, s); return 0; }
      if (phys == FECX_J2) { launch_modes<FECX_TET10, 3, FECX_J2>(mode, a, s); return 0; }
    }
  }
  return -1;
}

}  // namespace fecx
