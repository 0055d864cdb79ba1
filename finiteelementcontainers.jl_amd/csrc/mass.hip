// Mass-type element kernels (phase 1 of the mass assemblies): the consistent mass blocks, the row-sum lumped mass,
// the diagonal of the consistent mass and the mass action. They depend on the shape functions and the isoparametric
// map only, so they skip the field gather and the material: per element, the quadrature lanes compute
// w_q = rho det(J_q) W_q and N_a(xi_q), and the node lanes form m_ab = sum_q w_q N_a N_b.
//
// Reference: the test physics' mass callbacks M_el = rho JxW kron(N N^T, I_NF) and m_el = rho JxW N_a
// (test/mechanics/TestMechanicsCommon.jl:52-124, rho = props[1]; test/poisson/TestPoissonCommon.jl:18-41, rho = 1),
// assembled by assemble_mass! / assemble_lumped_mass! / assemble_diagonal! / assemble_matrix_action!
// (src/assemblers/Matrix.jl:1-10, LumpedMass.jl:29-58, Diagonal.jl:17-68, MatrixAction.jl:154-238).
// The outputs use the element kernels' storage (element-private FP32 Rel / symmetric NF x NF pair blocks in Kel),
// so the FP64 reductions of reduce.hip sum them into the field or the CSR values unchanged.
#include "fem_device.h"

namespace fecx {

template <int KIND, int NF, int WHAT>
struct MassLDS {
  using T = ET<KIND>;
  float x[T::NNPE][T::ND];
  float v[WHAT == MASS_ACTION ? T::NNPE : 1][NF];
  struct QP {
    float w;
    float N[T::NNPE];
  };
  static constexpr int KW = WHAT == MASS_MATRIX ? T::NPAIRS * NF * NF : 1;
};

constexpr int MASS_THREADS = 256;

// The work-group's quadrature data and, over it once consumed, its element-matrix blocks, contiguous per wavefront
// (the element kernels' layout: 16-byte copies to HBM, per-element qp blocks padded to 8 (mod 16) words).
template <class LDS, int EPW, int NQ>
struct MassWork {
  union {
    QPBlock<typename LDS::QP, NQ, qp_pad<typename LDS::QP, NQ>()> qp[EPW];
    float K[EPW * LDS::KW];
  };
};

template <int KIND, int NF, int WHAT>
__global__ __launch_bounds__(256) void mass_kernel(ElementArgs args, float rho) {
  using T = ET<KIND>;
  using LDS = MassLDS<KIND, NF, WHAT>;
  constexpr int ND = T::ND, NNPE = T::NNPE, NQ = T::NQ, G = T::G, EPW = MASS_THREADS / G, NDOF = NNPE * NF;
  constexpr int MAXP = T::MAXP;
  __shared__ LDS lds[EPW];
  __shared__ __attribute__((aligned(16))) MassWork<LDS, EPW, NQ> work;
  constexpr int KW = LDS::KW;

  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const int64_t e = args.ebeg + (int64_t)blockIdx.x * EPW + grp;  // element storage position
  const bool valid = e < args.ne;
  auto& L = lds[grp];
  bool bad = false;

  // ---------------- gather (lane = local node): element-local coordinates x_a - x_0 from the hi/lo split
  {
    float hi[ND] = {}, lo[ND] = {}, vv[NF] = {};
    if (valid && lane < NNPE) {
      const int64_t node = args.conn[e * NNPE + lane];
#pragma unroll
      for (int d = 0; d < ND; ++d) { hi[d] = args.xhi[node * ND + d]; lo[d] = args.xlo[node * ND + d]; }
      if constexpr (WHAT == MASS_ACTION) {
#pragma unroll
        for (int i = 0; i < NF; ++i) vv[i] = (float)args.V[node * NF + i];
      }
    }
    const int src0 = (grp * G) % 64;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const float h0 = __shfl(hi[d], src0), l0 = __shfl(lo[d], src0);
      if (lane < NNPE) L.x[lane][d] = (hi[d] - h0) + (lo[d] - l0);
    }
    if constexpr (WHAT == MASS_ACTION) {
      if (lane < NNPE)
#pragma unroll
        for (int i = 0; i < NF; ++i) L.v[lane][i] = vv[i];
    }
  }
  __syncthreads();

  // ---------------- quadrature points (lane = q): weight and shape values
  if (valid && lane < NQ) {
    const int q = lane;
    float xi[ND];
    qpoint<KIND>(q, xi);
    float N[NNPE], dN[NNPE][ND];
#pragma unroll
    for (int a = 0; a < NNPE; ++a) shape_node<KIND>(a, xi, N[a], dN[a]);
    float J[ND][ND] = {};
#pragma unroll
    for (int a = 0; a < NNPE; ++a)
#pragma unroll
      for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) J[i][j] += L.x[a][i] * dN[a][j];
    float det;
    if constexpr (ND == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    } else {
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
    if (!(det > 0.0f)) bad = true;
    work.qp[grp].q[q].w = rho * det * qweight<KIND>();
#pragma unroll
    for (int a = 0; a < NNPE; ++a) work.qp[grp].q[q].N[a] = N[a];
  }
  __syncthreads();

  // ---------------- node rows (lane = a)
  float kreg[WHAT == MASS_MATRIX ? MAXP : 1];
  if (valid && lane < NNPE) {
    const int a = lane;
    if constexpr (WHAT == MASS_LUMPED || WHAT == MASS_DIAG) {
      float s = 0.0f;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const float na = work.qp[grp].q[q].N[a];
        s = fmaf(work.qp[grp].q[q].w, WHAT == MASS_DIAG ? na * na : na, s);
      }
#pragma unroll
      for (int i = 0; i < NF; ++i) args.Rel[e * NDOF + a * NF + i] = s;
    } else if constexpr (WHAT == MASS_ACTION) {
      float out[NF] = {};
      for (int b = 0; b < NNPE; ++b) {
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < NQ; ++q) m = fmaf(work.qp[grp].q[q].w * work.qp[grp].q[q].N[a], work.qp[grp].q[q].N[b], m);
#pragma unroll
        for (int i = 0; i < NF; ++i) out[i] = fmaf(m, L.v[b][i], out[i]);
      }
#pragma unroll
      for (int i = 0; i < NF; ++i) args.Rel[e * NDOF + a * NF + i] = out[i];
    } else {
      const int cnt = lane_cnt(NNPE, a);
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        float m = 0.0f;
        if (p < cnt) {
          const int b = a + p < NNPE ? a + p : a + p - NNPE;
#pragma unroll
          for (int q = 0; q < NQ; ++q) m = fmaf(work.qp[grp].q[q].w * work.qp[grp].q[q].N[a], work.qp[grp].q[q].N[b], m);
        }
        kreg[p] = m;
      }
    }
  }
  if constexpr (WHAT == MASS_MATRIX) {
    // stage the blocks m_ab I in LDS over the consumed quadrature data, then copy each wavefront's contiguous range
    // of element storage with consecutive lanes on consecutive words (the element kernels' store)
    __syncthreads();
    if (valid && lane < NNPE) {
      const int cnt = lane_cnt(NNPE, lane);
      float* dst = work.K + grp * KW;
#pragma unroll
      for (int p = 0; p < MAXP; ++p)
        if (p < cnt) {
          const int b = lane + p < NNPE ? lane + p : lane + p - NNPE;
          float* blk = dst + pair_pos(NNPE, lane, b, pair_lane_off(NNPE, lane) + p) * (NF * NF);
#pragma unroll
          for (int c = 0; c < NF * NF; ++c) blk[c] = (c % (NF + 1) == 0) ? kreg[p] : 0.0f;
        }
    }
    __syncthreads();
    constexpr int EW = T::NPAIRS * NF * NF;
    constexpr int EPWAVE = 64 / G;
    const int wl = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int64_t e0 = args.ebeg + (int64_t)blockIdx.x * EPW + wv * EPWAVE;
    const int64_t nvalid = args.ne - e0 < EPWAVE ? args.ne - e0 : EPWAVE;
    static_assert(EW == KW && (EPWAVE * EW) % 4 == 0, "a wavefront's blocks start 16-byte aligned in LDS");
    if (nvalid > 0) {
      float* out = args.Kel + e0 * EW;
      const float* src = work.K + wv * EPWAVE * EW;
      const int total = (int)nvalid * EW;
      int done = 0;
      if (((e0 * EW) & 3) == 0) {  // 16-byte aligned in HBM too
        done = total & ~3;
        for (int i = 4 * wl; i < done; i += 256)
          __builtin_nontemporal_store(*reinterpret_cast<const f32x4*>(src + i), reinterpret_cast<f32x4*>(out + i));
      }
      for (int idx = done + wl; idx < total; idx += 64) out[idx] = src[idx];
    }
  }
  if (bad) atomicOr(args.flag, 1);
}

template <int KIND, int NF>
static void launch_mass_kind(int what, const ElementArgs& a, float rho, hipStream_t s) {
  using T = ET<KIND>;
  constexpr int EPW = MASS_THREADS / T::G;
  if (a.ne <= a.ebeg) return;
  const unsigned grid = (unsigned)((a.ne - a.ebeg + EPW - 1) / EPW);
  switch (what) {
    case MASS_MATRIX: mass_kernel<KIND, NF, MASS_MATRIX><<<grid, MASS_THREADS, 0, s>>>(a, rho); break;
    case MASS_LUMPED: mass_kernel<KIND, NF, MASS_LUMPED><<<grid, MASS_THREADS, 0, s>>>(a, rho); break;
    case MASS_DIAG: mass_kernel<KIND, NF, MASS_DIAG><<<grid, MASS_THREADS, 0, s>>>(a, rho); break;
    case MASS_ACTION: mass_kernel<KIND, NF, MASS_ACTION><<<grid, MASS_THREADS, 0, s>>>(a, rho); break;
    default: break;
  }
}

int launch_mass(int kind, int nf, int what, const ElementArgs& a, float rho, hipStream_t s) {
  if (what < MASS_MATRIX || what > MASS_ACTION) return -1;
  if (nf == 1) {
    if (kind == FECX_QUAD4) { launch_mass_kind<FECX_QUAD4, 1>(what, a, rho, s); return 0; }
    if (kind == FECX_HEX8) { launch_mass_kind<FECX_HEX8, 1>(what, a, rho, s); return 0; }
    if (kind == FECX_TET10) { launch_mass_kind<FECX_TET10, 1>(what, a, rho, s); return 0; }
  }
  if (nf == 2 && kind == FECX_QUAD4) { launch_mass_kind<FECX_QUAD4, 2>(what, a, rho, s); return 0; }
  if (nf == 3) {
    if (kind == FECX_HEX8) { launch_mass_kind<FECX_HEX8, 3>(what, a, rho, s); return 0; }
    if (kind == FECX_TET10) { launch_mass_kind<FECX_TET10, 3>(what, a, rho, s); return 0; }
  }
  return -1;
}

}  // namespace fecx
