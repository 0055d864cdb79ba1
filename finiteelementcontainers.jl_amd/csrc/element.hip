// Element kernels (phase 1 of assembly): FP32 element math, element-private FP32 outputs.
//
// Reference: the generic element loop _assemble_block! (src/assemblers/Assemblers.jl:402-432) runs one
// work-item per element and stores K_el into element-private COO slots (Assemblers.jl:109-124). Here a
// 256-thread work-group processes 256/G elements with G lanes per element:
//   gather   lane a  (a < nnpe) : connectivity, coordinates and nodal values of local node a; element-local
//                                 differences x_a - x_0 (hi/lo FP32 coordinate split) and u_a - u_0
//   qp       lane q  (q < nq)   : isoparametric map, shape gradients, field gradient and material point of
//                                 quadrature point q: stress and a compact form of the tangent, staged in LDS
//                                 (tet10: lanes 4q .. 4q+3, each with every 4th node)
//   node     lane a  (a < nnpe) : residual / action rows of node a, and the node-pair blocks K_ab it owns
//                                 (symmetric storage, one NF x NF block per unordered pair)
// Each material contracts its tangent with the shape gradients in closed form (K_ab = sum_q w g_a . A . g_b
// written through a few per-node vectors), so the LDS holds O(10) floats of tangent data per quadrature point
// instead of the 45 of a dense symmetric 9 x 9 tangent.
// Outputs are FP32 element-private slots in element storage order; the FP64 sums into the residual field and
// the CSR values happen in the reduction kernel (reduce.hip), through the precomputed inverse COO->CSR map.
#include "fem_device.h"

namespace fecx {

// ------------------------------------------------------------------ the kernel
// The element's coordinate and field differences to its local node 0 (in the element record, or, IN_WORK, elsewhere)
template <class RT, int NNPE, int ND, int NF, bool IN_WORK>
struct NodeDiffs {
  RT x[NNPE][ND];
  RT u[NNPE][NF];
};
template <class RT, int NNPE, int ND, int NF>
struct NodeDiffs<RT, NNPE, ND, NF, true> {};

// tet10 modes that stage element blocks keep the node differences in the work area (see GW below)
template <int KIND, int MODE>
constexpr bool diffs_in_work() {
  return KIND == FECX_TET10 && (MODE & (MODE_K | MODE_ACTION_ASM)) != 0;
}

template <int KIND, int NF, int PHYS, int MODE, bool G64 = false>
struct ElemLDS : NodeDiffs<std::conditional_t<G64, double, float>, ET<KIND>::NNPE, ET<KIND>::ND, NF,
                           diffs_in_work<KIND, MODE>()> {
  using T = ET<KIND>;
  static constexpr bool DO_R = (MODE & MODE_R) != 0, DO_K = (MODE & MODE_K) != 0;
  static constexpr bool DO_MF = (MODE & MODE_ACTION_MF) != 0, DO_ASM = (MODE & MODE_ACTION_ASM) != 0;
  static constexpr bool DO_D = (MODE & MODE_DIAG) != 0;
  static constexpr bool NEED_C = DO_K || DO_MF || DO_ASM || DO_D;
  static constexpr int D = NF * T::ND;
  // Compact quadrature data (hex8 vector tangents): per node only the shape gradient g_a, plus the point's tangent data
  // C once; the node lanes form the per-node vectors prep(C, g) themselves. The neo-Hookean and J2 blocks then stage
  // 45 / 44 instead of 60 floats per point: under 53 KB per work-group, three work-groups per CU instead of two.
  static constexpr bool GS = KIND == FECX_HEX8 && NF == 3 && (DO_K || DO_ASM || DO_D) && (Mat<PHYS>::NV > T::ND);
  // Poisson hex8 residual formed in the quadrature phase: each point's lane computes its contribution to all 8 node
  // rows and three DPP steps sum them over the element's 8 points, so the flux and source need no LDS (30 KB instead
  // of 36 KB per work-group: five work-groups per CU instead of four).
  // The element blocks need only sqrt(w) g_a per point and node (24 floats per point); every Poisson hex8 mode that
  // forms them stages that (SC), so fused and separate calls give the same bits.
  // The matrix-free action is summed the same way (QM) and stages nothing. The other hex8 materials sum their residual
  // and action this way too, in every mode (fused and separate calls give the same bits); where no element block is
  // formed they stage nothing (NOQP), so LDS no longer limits their occupancy.
  static constexpr bool HEX = KIND == FECX_HEX8 && (MODE & MODE_QPGRAD) == 0;
  static constexpr bool SC = HEX && PHYS == FECX_POISSON;
  static constexpr bool QR = HEX && DO_R, QM = HEX && DO_MF;
  static constexpr bool NOQP = HEX && !(DO_K || DO_ASM || DO_D);  // no node-phase consumer of quadrature data
  // G64 (tet10, fecx_set_fp64_geometry): the element geometry, field gradient and stress are evaluated in FP64 (from
  // FP64 coordinate and field differences); their FP32 roundings feed the tangent, the residual contraction and the
  // stored rows. tet10 modes that stage element blocks: the differences (Geo) live in the wavefront's work area beside
  // the quadrature data (GW), in space that the element blocks take only once both are consumed.
  static_assert(!G64 || KIND == FECX_TET10, "FP64 element geometry for tet10 only");
  using RT = std::conditional_t<G64, double, float>;
  static constexpr bool GW = diffs_in_work<KIND, MODE>();
  using Geo = NodeDiffs<RT, T::NNPE, T::ND, NF, false>;
  float x0[T::ND];
  float v[(DO_MF || DO_ASM) ? T::NNPE : 1][NF];
  struct QPFull {
    float v[T::NNPE][(DO_K || DO_ASM || DO_D) && !GS ? Mat<PHYS>::NV : T::ND];  // v[a][0..ND) = shape gradient g_a
    float w;
    float P[(DO_R && !QR) || (DO_MF && !QM) ? D : 1];
    float src[(DO_R && PHYS == FECX_POISSON) ? 1 : 1];
    union {
      float S[Mat<PHYS>::NS > 0 && !GS ? Mat<PHYS>::NS : 1];  // side scalars of the tangent (C[S0 .. S0+NS))
      float C[GS ? Mat<PHYS>::NC : 1];                          // compact form: the whole tangent data
    };
  };
  struct QPScaled {
    float v[T::NNPE][T::ND];  // sqrt(w) g_a
  };
  using QP = std::conditional_t<SC || NOQP, QPScaled, QPFull>;
  static constexpr int KW = (DO_K || DO_ASM) ? T::NPAIRS * NF * NF : 1;  // staged element-matrix words
};

constexpr int ELEM_THREADS = 256;  // four wavefronts, 256 / G elements per work-group

// Quadrature data of the work-group's elements, and over it, once consumed, their element-matrix blocks: those of one
// wavefront's elements are contiguous, in element storage order, so the wavefront copies them to HBM as 16-byte
// vectors (one 1 KB store per instruction) with no per-word element/offset arithmetic.
// KS: words between two elements' blocks (LDS::KW, or more where a consumer wants another bank spread).
template <class LDS, int EPW, int NQ, int KS = LDS::KW>
struct ElemWork {
  using Block = QPBlock<typename LDS::QP, NQ, qp_pad<typename LDS::QP, NQ>()>;
  static constexpr bool WAVE_LOCAL = false;  // element blocks overlap other wavefronts' quadrature data
  static_assert(!LDS::GW, "element differences in the element records");
  union {
    Block qp[EPW];
    float K[EPW * KS];
  };
  __device__ Block& qpb(int e) { return qp[e]; }
  __device__ float* kb(int e) { return K + e * KS; }
};

// The same per wavefront: each wavefront's quadrature data and, over them, its elements' blocks (contiguous, LDS::KW
// words apart) in a region of their own, so that every LDS word an element's lanes write or read between the phases
// belongs to their own wavefront: wavefront-local ordering replaces the work-group barriers (element_kernel).
template <class LDS, int EPW, int NQ, int G>
struct WaveWork {
  using Block = QPBlock<typename LDS::QP, NQ, qp_pad<typename LDS::QP, NQ>()>;
  static constexpr bool WAVE_LOCAL = true;
  static constexpr int EPWAVE = 64 / G;
  static_assert(EPW % EPWAVE == 0, "whole wavefronts of elements");
  struct QPOnly {
    Block qp[EPWAVE];
  };
  struct QPGeo {
    Block qp[EPWAVE];
    typename LDS::Geo geo[EPWAVE];
  };
  union Wave {
    std::conditional_t<LDS::GW, QPGeo, QPOnly> a;
    float K[EPWAVE * LDS::KW];
  };
  Wave wv[EPW / EPWAVE];
  static_assert(sizeof(Wave) % 16 == 0, "every wavefront's blocks start 16-byte aligned");
  __device__ Block& qpb(int e) { return wv[e / EPWAVE].a.qp[e % EPWAVE]; }
  __device__ float* kb(int e) { return wv[e / EPWAVE].K + (e % EPWAVE) * LDS::KW; }
  __device__ typename LDS::Geo& geo(int e) { return wv[e / EPWAVE].a.geo[e % EPWAVE]; }
};

// Ordering between the phases of element_eval: a work-group barrier, or (WAVE_LOCAL work areas) the program order of
// one wavefront's LDS operations, which complete in order, kept by a wavefront-scope fence.
template <class WORK>
__device__ __forceinline__ void phase_sync() {
  if constexpr (WORK::WAVE_LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// A lane permutation within rows of 16 lanes (quad_perm, row_half_mirror: every lane has a source). bound_ctrl set,
// which changes nothing for these permutations, lets the compiler fold the move into the add that consumes it
// (v_add_f32_dpp).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov_f64(double x) {  // the same permutation of both 32-bit halves
  const int64_t b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

// Reduce-scatter over the 8 quadrature-point lanes of a hex8 element. On entry each lane holds v[a * NF + i], its
// point's contribution to node a, component i; on exit lane a holds out[i], the sum of those over the 8 points. Three
// exchanges halve the nodes a lane keeps: with lane 7 - q (half-row mirror; lanes 0-3 keep nodes 0-3), then with
// q ^ 2 and q ^ 1 (quad permutations): 7 NF exchanges, where one all-reduce per value takes 24 NF.
template <int NF>
__device__ __forceinline__ void scatter_points8(const float (&v)[8 * NF], int q, float (&out)[NF]) {
  const bool h = (q & 4) != 0, b = (q & 2) != 0, c = (q & 1) != 0;
  float v1[4 * NF], v2[2 * NF];
#pragma unroll
  for (int k = 0; k < 4 * NF; ++k) v1[k] = (h ? v[4 * NF + k] : v[k]) + dpp_mov<0x141>(h ? v[k] : v[4 * NF + k]);
#pragma unroll
  for (int k = 0; k < 2 * NF; ++k) v2[k] = (b ? v1[2 * NF + k] : v1[k]) + dpp_mov<0x4E>(b ? v1[k] : v1[2 * NF + k]);
#pragma unroll
  for (int k = 0; k < NF; ++k) out[k] = (c ? v2[NF + k] : v2[k]) + dpp_mov<0xB1>(c ? v2[k] : v2[NF + k]);
}

// Element-matrix blocks formed by one lane of the node phase. Every lane forms the blocks K_cb of one centre node c
// (one evaluation of c's side terms per point). quad4 and hex8: c = lane, b = c + p for the lane's slots
// (pair_lane_off order). tet10 (16 lanes, 55 blocks): lanes 0-9 form b = c + 0..3; lanes 10-15 form the 15 blocks of
// offset 4 and 5 around c = lane - 10 as their blocks 1-3: K_c,c+4 and K_c,c+5 as stored, K_c,c+6 transposed (its
// slot holds K_c+6,c, equal by the tangent's major symmetry). The busiest lane forms 4 blocks instead of 6, and block 0
// of every lane that has one is its diagonal block K_cc.
template <int KIND>
struct PairJobs {
  static constexpr int NNPE = ET<KIND>::NNPE;
  static constexpr int LANES = KIND == FECX_TET10 ? 16 : NNPE;  // lanes that form blocks
  static constexpr int NP = KIND == FECX_TET10 ? 4 : ET<KIND>::MAXP;  // blocks per lane, at most
  __host__ __device__ static constexpr int center(int lane) { return lane < NNPE ? lane : lane - NNPE; }
  // block p of `lane`: false if the lane has none, else the other node b, the storage slot and whether the slot holds
  // the transpose of the computed K_cb
  __host__ __device__ static constexpr bool job(int lane, int p, int& b, int& slot, bool& tr) {
    const int c = center(lane);
    if (KIND != FECX_TET10 || lane < NNPE) {
      const int lim = KIND == FECX_TET10 ? 4 : lane_cnt(NNPE, c);
      b = c + p < NNPE ? c + p : c + p - NNPE;
      slot = pair_lane_off(NNPE, c) + p;
      tr = false;
      return p < lim;
    }
    // no block at p = 0, where lanes 0-9 form their diagonal blocks (a cheaper form that then runs alone)
    const int j = p - 1;
    const int d = j == 0 ? 4 : (j == 1 ? 6 : 5);
    const int o = j == 1 ? c + 6 : c;  // the node whose slots hold the block
    b = c + d < NNPE ? c + d : c + d - NNPE;
    slot = pair_lane_off(NNPE, o) + (j == 2 ? 5 : 4);
    tr = j == 1;
    return j == 0 ? c <= 5 : (j == 1 ? c <= 3 : (j == 2 ? c <= 4 : false));
  }
};

// tet10 block addressing, built at compile time from PairJobs, pair_ref and pair_pos, so that a lane loads its bytes
// once instead of evaluating their index arithmetic (~110 instructions per wavefront in the staging alone):
//   stage[lane]: byte p = job p of the lane's node phase: position of its block in the element record (bits 0-5),
//                the block is stored transposed (bit 6), the job exists (bit 7);
//   rows[a][b]:  position of the block of the pair {a, b} (bits 0-5), the block is K_ab as stored (bit 6).
struct Tet10Addr {
  uint32_t stage[16];
  uint8_t rows[10][12];
};
constexpr Tet10Addr make_tet10_addr() {
  Tet10Addr t{};
  using PJ = PairJobs<FECX_TET10>;
  for (int lane = 0; lane < 16; ++lane) {
    uint32_t w = 0;
    for (int p = 0; p < PJ::NP; ++p) {
      int b = 0, slot = 0;
      bool tr = false;
      if (PJ::job(lane, p, b, slot, tr))
        w |= (uint32_t)(pair_pos(10, PJ::center(lane), b, slot) | (tr ? 64 : 0) | 128) << (8 * p);
    }
    t.stage[lane] = w;
  }
  for (int a = 0; a < 10; ++a)
    for (int b = 0; b < 10; ++b) {
      const PairRef pr = pair_ref(10, a, b);
      t.rows[a][b] = (uint8_t)(pair_pos(10, a, b, pr.slot) | (pr.ab ? 64 : 0));
    }
  return t;
}
__constant__ Tet10Addr c_tet10_addr = make_tet10_addr();
constexpr int TET10_TR_JOB = 2;  // the one job of the tet10 lanes whose blocks are stored transposed
constexpr bool tet10_transposes_one_job() {
  const Tet10Addr t = make_tet10_addr();
  for (int lane = 0; lane < 16; ++lane)
    for (int p = 0; p < 4; ++p)
      if (p != TET10_TR_JOB && ((t.stage[lane] >> (8 * p)) & 64u)) return false;
  return true;
}
static_assert(tet10_transposes_one_job(), "tet10 staging: only job TET10_TR_JOB is stored transposed");

// One lane's gather inputs: its local node's coordinates (hi/lo FP32 split) and field values. Loaded apart from the
// element computation so that a caller can issue the loads early (the fused kernel fetches the next layer's while
// the current one is reduced).
template <int ND, int NF>
struct NodeIn {
  float hi[ND], lo[ND];
  double u[NF], v[NF];
};
template <int ND, int NF>
__device__ __forceinline__ NodeIn<ND, NF> load_node_in(const ElementArgs& args, int64_t node, bool load, bool with_v) {
  NodeIn<ND, NF> in;
#pragma unroll
  for (int d = 0; d < ND; ++d) { in.hi[d] = 0.0f; in.lo[d] = 0.0f; }
#pragma unroll
  for (int i = 0; i < NF; ++i) { in.u[i] = 0.0; in.v[i] = 0.0; }
  if (load) {
#pragma unroll
    for (int d = 0; d < ND; ++d) { in.hi[d] = args.xhi[node * ND + d]; in.lo[d] = args.xlo[node * ND + d]; }
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      in.u[i] = args.U[node * NF + i];
      if (with_v) in.v[i] = args.V[node * NF + i];
    }
  }
  return in;
}

// The element computation of one 256-thread work-group, G lanes per element (gather, quadrature points, node phase).
// Lane group grp = threadIdx.x / G evaluates the element at storage position e (when `valid`). Every element row
// (residual, action, diagonal) leaves through rel_out(word a * NF + i, value); the element blocks (DO_K) are left in
// the work area, element grp's at work.kb(grp) in its storage layout, when this returns (after a work-group barrier,
// or for a wavefront-local work area an ordering of the wavefront's own LDS operations). Used by element_kernel
// (two-phase assembly, WaveWork) and by the structured hex8 kernels (ElemWork; they reduce the blocks from LDS). `in` holds the lane's gather inputs (lanes below NNPE of a valid element; load_node_in), `sel`
// the element id that indexes the quadrature-point state; so_pre, if given, the old state of the lane's point (J2).
template <int KIND, int NF, int PHYS, int MODE, int KS = 0, class LDS, class WORK, class RelOut>
__device__ __forceinline__ void element_eval(const ElementArgs& args, int64_t e, int64_t sel, bool valid,
                                             const NodeIn<ET<KIND>::ND, NF>& in, LDS* lds, WORK& work,
                                             RelOut rel_out, bool& bad, int tid, const float* so_pre = nullptr) {
  using T = ET<KIND>;
  using M = Mat<PHYS>;
  constexpr int ND = T::ND, NNPE = T::NNPE, NQ = T::NQ, G = T::G;
  constexpr int D = NF * ND;
  constexpr bool DO_R = LDS::DO_R, DO_K = LDS::DO_K, DO_MF = LDS::DO_MF, DO_ASM = LDS::DO_ASM, DO_D = LDS::DO_D;
  constexpr bool DO_G = (MODE & MODE_QPGRAD) != 0;
  constexpr bool NEED_C = LDS::NEED_C, GS = LDS::GS, SC = LDS::SC, QR = LDS::QR, QM = LDS::QM, NOQP = LDS::NOQP;
  constexpr int KW = KS ? KS : LDS::KW;
  static_assert(KW >= LDS::KW, "element block stride below the element's words");
  using RT = typename LDS::RT;  // the scalar of the element geometry and stress chain

  const int lane = tid % G, grp = tid / G;
  auto& L = lds[grp];
  // node a's coordinate and field differences to local node 0
  auto Xn = [&](int a) -> RT* {
    if constexpr (LDS::GW) return work.geo(grp).x[a];
    else return L.x[a];
  };
  auto Un = [&](int a) -> RT* {
    if constexpr (LDS::GW) return work.geo(grp).u[a];
    else return L.u[a];
  };

  // The node lanes' view of quadrature point Q: its tangent data C (the side scalars C[S0 .. S0+NS), or all of it in
  // the compact form), and node b's vector vb (staged, or formed here from g_b by the same prep as the staged form).
  auto load_c = [&](const auto& Q, float* C) {
    if constexpr (GS) {
#pragma unroll
      for (int c = 0; c < M::NC; ++c) C[c] = Q.C[c];
    } else if constexpr (M::NS > 0) {
#pragma unroll
      for (int c = 0; c < M::NS; ++c) C[M::S0 + c] = Q.S[c];
    }
  };
  auto qp_weight = [](const auto& Q) {  // the scaled record holds sqrt(w) g: weight 1 in the stiffness
    if constexpr (SC) return 1.0f;
    else return Q.w;
  };
  auto node_vec = [&](const auto& Q, const float* C, int b, float* vb) {
    if constexpr (GS) {
      float gb[ND];
#pragma unroll
      for (int j = 0; j < ND; ++j) gb[j] = Q.v[b][j];
      M::template prep<ND>(args.mp, C, gb, vb);
    } else {
#pragma unroll
      for (int c = 0; c < M::NV; ++c) vb[c] = Q.v[b][c];
    }
  };

  // ---------------- gather (lane = local node)
  {
    const float* hi = in.hi;
    const float* lo = in.lo;
    const double* uu = in.u;
    const double* vv = in.v;
    const int src0 = (grp * G) % 64;  // lane (within the wavefront) of local node 0
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const float h0 = __shfl(hi[d], src0), l0 = __shfl(lo[d], src0);
      if constexpr (std::is_same_v<RT, double>) {
        if (lane < NNPE) Xn(lane)[d] = ((double)hi[d] - (double)h0) + ((double)lo[d] - (double)l0);
      } else {
        if (lane < NNPE) Xn(lane)[d] = (hi[d] - h0) + (lo[d] - l0);
      }
      if (lane == 0) L.x0[d] = h0 + l0;
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      const double u0 = __shfl(uu[i], src0);
      if (lane < NNPE) Un(lane)[i] = (RT)(uu[i] - u0);
      if constexpr (DO_MF) {
        const double v0 = __shfl(vv[i], src0);
        if (lane < NNPE) L.v[lane][i] = (float)(vv[i] - v0);
      }
      if constexpr (DO_ASM) {
        if (lane < NNPE) L.v[lane][i] = (float)vv[i];
      }
    }
  }
  phase_sync<WORK>();

  // ---------------- quadrature points: S lanes per point, lane S q + r handling the nodes r, r + S, r + 2S, ...
  // (tet10: 4 points on 16 lanes, S = 4; quad4/hex8: S = 1). The node sums (Jacobian, field gradient, source point,
  // action gradient) are completed across the point's 4 lanes by two DPP quad permutations (xor 1, then xor 2: VALU
  // operations, no LDS round trip), which leave the same bits on every lane of the point, so each lane evaluates the
  // material itself and writes the quadrature data of its own nodes.
  constexpr int S = KIND == FECX_TET10 ? G / NQ : 1;
  constexpr int NMY = (NNPE + S - 1) / S;
  static_assert(NQ * S <= G && (S == 1 || S == 4), "the lanes of an element cover its quadrature points in quads");
  auto lane_sum = [](auto& v) {
    if constexpr (S == 4) {
      if constexpr (std::is_same_v<std::decay_t<decltype(v)>, double>) {
        v += dpp_mov_f64<0xB1>(v);  // quad_perm 1,0,3,2
        v += dpp_mov_f64<0x4E>(v);  // quad_perm 2,3,0,1
      } else {
        v += dpp_mov<0xB1>(v);  // quad_perm 1,0,3,2
        v += dpp_mov<0x4E>(v);  // quad_perm 2,3,0,1
      }
    }
  };
  if (valid && lane < NQ * S) {
    const int q = lane / S, r = lane % S;
    auto mine = [&](int k) { return S == 1 || r + S * k < NNPE; };
    RT xi[ND];
    qpoint<KIND>(q, xi);
    RT N[NMY], dN[NMY][ND];
#pragma unroll
    for (int k = 0; k < NMY; ++k) {
      N[k] = RT(0);
#pragma unroll
      for (int j = 0; j < ND; ++j) dN[k][j] = RT(0);
      if (mine(k)) shape_node<KIND>(r + S * k, xi, N[k], dN[k]);
    }
    RT J[ND][ND] = {};
#pragma unroll
    for (int k = 0; k < NMY; ++k)
      if (mine(k))
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
          for (int j = 0; j < ND; ++j) J[i][j] += Xn(r + S * k)[i] * dN[k][j];
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int j = 0; j < ND; ++j) lane_sum(J[i][j]);
    RT Ji[ND][ND], det;
    if constexpr (ND == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const RT id = RT(1) / det;
      Ji[0][0] = J[1][1] * id; Ji[0][1] = -J[0][1] * id; Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
    } else {
      const RT c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
               c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const RT id = RT(1) / det;
      Ji[0][0] = c00 * id; Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[1][0] = c01 * id; Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[2][0] = c02 * id; Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
    if (!(det > RT(0))) bad = true;
    auto& Q = work.qpb(grp).q[q];
    RT gr[NMY][ND];
#pragma unroll
    for (int k = 0; k < NMY; ++k)
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        RT t = 0;
#pragma unroll
        for (int i = 0; i < ND; ++i) t += dN[k][i] * Ji[i][j];
        gr[k][j] = t;
      }
    const RT wr = det * qweight<KIND, RT>();
    RT Hr[D] = {};
#pragma unroll
    for (int k = 0; k < NMY; ++k)
      if (mine(k))
#pragma unroll
        for (int i = 0; i < NF; ++i)
#pragma unroll
          for (int j = 0; j < ND; ++j) Hr[i * ND + j] += Un(r + S * k)[i] * gr[k][j];
#pragma unroll
    for (int c = 0; c < D; ++c) lane_sum(Hr[c]);
    // FP32 shape gradients, weight and field gradient from here on (the same values where RT is float)
    float g[NMY][ND], H[D];
#pragma unroll
    for (int k = 0; k < NMY; ++k)
#pragma unroll
      for (int j = 0; j < ND; ++j) g[k][j] = (float)gr[k][j];
#pragma unroll
    for (int c = 0; c < D; ++c) H[c] = (float)Hr[c];
    const float w = (float)wr;
    if constexpr (!SC && !NOQP) {
      if (r == 0) Q.w = w;
    }
    if constexpr (!(DO_K || DO_ASM || DO_D) && !SC && !NOQP) {  // modes without element blocks: g_a as it is, stored
#pragma unroll                                                  // before the material so that it need not stay live
      for (int k = 0; k < NMY; ++k)
        if (mine(k))
#pragma unroll
          for (int j = 0; j < ND; ++j) Q.v[r + S * k][j] = g[k][j];
    }
    if constexpr (DO_G) {  // quadrature-point gradients only, element storage order
      if (r == 0) {
        float* o = args.hq + ((int64_t)e * NQ + q) * D;
#pragma unroll
        for (int c = 0; c < D; ++c) o[c] = H[c];
      }
    } else {
    float P[D], C[M::NC];
    if constexpr (M::FROM_QP) {  // stress and tangent data from the batched (surrogate) material pass
      const float* qm = args.qpmat + ((int64_t)e * NQ + q) * (D + M::NC);
#pragma unroll
      for (int c = 0; c < D; ++c) P[c] = qm[c];
#pragma unroll
      for (int c = 0; c < M::NC; ++c) C[c] = qm[D + c];
    } else {
      float so[7] = {}, sn[7];
      if constexpr (PHYS == FECX_J2) {  // the point's old state: loaded here, or already by the caller (so_pre)
        const int64_t sidx = (sel * NQ + q) * 7;
#pragma unroll
        for (int st = 0; st < 7; ++st) so[st] = so_pre ? so_pre[st] : args.st_old[sidx + st];
      }
      RT Pr[D];
      M::template eval<ND>(args.mp, Hr, so, sn, Pr, C, bad);
#pragma unroll
      for (int c = 0; c < D; ++c) P[c] = (float)Pr[c];
      if constexpr (PHYS == FECX_J2) {
        if (DO_R && args.st_new && r == 0) {
          const int64_t sidx = (sel * NQ + q) * 7;
#pragma unroll
          for (int st = 0; st < 7; ++st) args.st_new[sidx + st] = sn[st];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (!(P[c] == P[c])) bad = true;
    if constexpr (DO_R) {
      if constexpr (!QR) {
        if (r == 0) {
#pragma unroll
          for (int c = 0; c < D; ++c) Q.P[c] = P[c];
        }
      }
      float f = 0.0f;  // Poisson source at the point
      if constexpr (PHYS == FECX_POISSON) {
        f = args.mp.p[1];
        if (args.mp.source_kind == 1) {  // the same for every lane of the launch
          f *= (float)ND * 9.8696044010893586f;
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            RT t = 0;
#pragma unroll
            for (int k = 0; k < NMY; ++k)
              if (mine(k)) t += N[k] * Xn(r + S * k)[d];
            lane_sum(t);
            f *= sinpif(L.x0[d] + (float)t);
          }
        }
        if constexpr (!QR) {
          if (r == 0) Q.src[0] = f;
        }
      }
      if constexpr (QR) {
        // R_a = sum_q w (P . g_a - N_a f) for the 8 nodes of this point, reduce-scattered over the element's 8
        // points: lane a keeps node a
        static_assert(NNPE == 8, "hex8 only");
        float wP[D], t[NNPE * NF], ra[NF];
#pragma unroll
        for (int c = 0; c < D; ++c) wP[c] = w * P[c];
        const float wf = w * f;
#pragma unroll
        for (int a = 0; a < NNPE; ++a)
#pragma unroll
          for (int i = 0; i < NF; ++i) {
            float s = 0;
#pragma unroll
            for (int j = 0; j < ND; ++j) s += wP[i * ND + j] * g[a][j];
            if constexpr (PHYS == FECX_POISSON) s -= N[a] * wf;
            t[a * NF + i] = s;
          }
        scatter_points8<NF>(t, q, ra);
#pragma unroll
        for (int i = 0; i < NF; ++i) rel_out(q * NF + i, ra[i]);
      }
    }
    if constexpr (GS) {
#pragma unroll
      for (int c = 0; c < M::NC; ++c) Q.C[c] = C[c];
#pragma unroll
      for (int a = 0; a < NNPE; ++a)
#pragma unroll
        for (int j = 0; j < ND; ++j) Q.v[a][j] = g[a][j];
    } else if constexpr (DO_K || DO_ASM || DO_D) {
      if constexpr (M::NS > 0) {
        if (r == 0) {
#pragma unroll
          for (int c = 0; c < M::NS; ++c) Q.S[c] = C[M::S0 + c];
        }
      }
      if constexpr (SC) {
        const float sw = sqrtf(w);  // K_ab = sum_q (sqrt(w) g_a) . (sqrt(w) g_b)
#pragma unroll
        for (int a = 0; a < NNPE; ++a)
#pragma unroll
          for (int j = 0; j < ND; ++j) Q.v[a][j] = sw * g[a][j];
      } else {
#pragma unroll
        for (int k = 0; k < NMY; ++k)
          if (mine(k)) {
            float va[M::NV];
            M::template prep<ND>(args.mp, C, g[k], va);
#pragma unroll
            for (int c = 0; c < M::NV; ++c) Q.v[r + S * k][c] = va[c];
          }
      }
    }
    if constexpr (DO_MF) {
      float Gv[D] = {};
#pragma unroll
      for (int k = 0; k < NMY; ++k)
        if (mine(k))
#pragma unroll
          for (int i = 0; i < NF; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) Gv[i * ND + j] += L.v[r + S * k][i] * g[k][j];
#pragma unroll
      for (int c = 0; c < D; ++c) lane_sum(Gv[c]);
      float out[D];
      M::template dP<ND>(args.mp, C, Gv, out);
      if constexpr (QM) {  // (K v)_a = sum_q w (A : Gv) . g_a, reduce-scattered over the 8 points like the residual
        float t[NNPE * NF], ra[NF];
#pragma unroll
        for (int c = 0; c < D; ++c) out[c] *= w;
#pragma unroll
        for (int a = 0; a < NNPE; ++a)
#pragma unroll
          for (int i = 0; i < NF; ++i) {
            float s = 0;
#pragma unroll
            for (int j = 0; j < ND; ++j) s += out[i * ND + j] * g[a][j];
            t[a * NF + i] = s;
          }
        scatter_points8<NF>(t, q, ra);
#pragma unroll
        for (int i = 0; i < NF; ++i) rel_out(q * NF + i, ra[i]);
      } else {
        if (r == 0) {
#pragma unroll
          for (int c = 0; c < D; ++c) Q.P[c] = out[c];
        }
      }
    }
    }  // !DO_G
  }
  phase_sync<WORK>();

  // ---------------- node rows (lane = a)
  using PJ = PairJobs<KIND>;
  constexpr int MAXP = PJ::NP;
  float kreg[(DO_K || DO_ASM) ? MAXP : 1][NF * NF];
  // tet10: the lane's block-staging addresses, loaded here so that the load is long done when the blocks are staged
  uint32_t stage_addr = 0;
  if constexpr (KIND == FECX_TET10 && (DO_K || DO_ASM)) stage_addr = c_tet10_addr.stage[lane];
  if (valid && lane < NNPE) {
    const int a = lane;
    if constexpr ((DO_R || DO_MF) && !QR && !QM) {
      float out[NF] = {};
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const auto& Q = work.qpb(grp).q[q];
        float Na = 0.0f, dNa[ND];
        if constexpr (DO_R && PHYS == FECX_POISSON) {
          float xi[ND];
          qpoint<KIND>(q, xi);
          shape_node<KIND>(a, xi, Na, dNa);
        }
#pragma unroll
        for (int i = 0; i < NF; ++i) {
          float s = 0;
#pragma unroll
          for (int j = 0; j < ND; ++j) s += Q.P[i * ND + j] * Q.v[a][j];
          if constexpr (DO_R && PHYS == FECX_POISSON) s -= Na * Q.src[i];
          out[i] += Q.w * s;
        }
      }
#pragma unroll
      for (int i = 0; i < NF; ++i) rel_out(a * NF + i, out[i]);
    }
  }
  // block 0 of every lane that has one is its diagonal block K_aa (PairJobs), formed by the material's cheaper
  // symmetric form
  constexpr bool DIAG0 = M::DIAG;
  if constexpr (DO_K || DO_ASM) {
    if (valid && lane < PJ::LANES) {
      const int a = PJ::center(lane);
      float acc[MAXP][NF * NF + M::KX];  // block words, then the material's extra accumulator (finish)
#pragma unroll
      for (int p = 0; p < MAXP; ++p)
#pragma unroll
        for (int c = 0; c < NF * NF + M::KX; ++c) acc[p][c] = 0.0f;
#pragma unroll 2
      for (int q = 0; q < NQ; ++q) {
        const auto& Q = work.qpb(grp).q[q];
        float C[M::NC];
        float va[M::NV], sa[M::NA];
        load_c(Q, C);
        node_vec(Q, C, a, va);
        M::template side<ND>(args.mp, C, qp_weight(Q), va, sa);
#pragma unroll
        for (int p = 0; p < MAXP; ++p) {
          int b, slot;
          bool tr;
          if (PJ::job(lane, p, b, slot, tr)) {
            if constexpr (DIAG0) {
              if (p == 0) {
                M::template pair_diag<ND>(sa, va, acc[0]);
                continue;
              }
            }
            float vb[M::NV];
            node_vec(Q, C, b, vb);
            M::template pair<ND>(sa, vb, acc[p]);
          }
        }
      }
#pragma unroll
      for (int p = 0; p < MAXP; ++p) M::template finish<ND>(acc[p]);
      if constexpr (NF > 1) {
        // the diagonal block K_aa (block 0 of lanes below NNPE) of these symmetric tangents is symmetric up to the
        // FP32 rounding of its two FMA orders; average it with its transpose (or, formed as its upper triangle by
        // pair_diag, mirror it) so that the assembled Jacobian is symmetric bit for bit
        if (lane < NNPE) {
#pragma unroll
          for (int i = 0; i < NF; ++i)
#pragma unroll
            for (int k = i + 1; k < NF; ++k) {
              const float m = DIAG0 ? acc[0][i * NF + k] : 0.5f * (acc[0][i * NF + k] + acc[0][k * NF + i]);
              acc[0][i * NF + k] = m;
              acc[0][k * NF + i] = m;
            }
        }
      }
      if constexpr (DO_K || DO_ASM) {
        // every lane of the work-group has finished reading the quadrature data of its element only after the
        // barrier below, so the blocks are kept in registers until then
#pragma unroll
        for (int p = 0; p < MAXP; ++p)
#pragma unroll
          for (int c = 0; c < NF * NF; ++c) kreg[p][c] = acc[p][c];
      }
    }
  }
  // the lane's blocks into the element's LDS block storage, each at its slot (transposed where the slot holds K_bc)
  auto stage_blocks = [&](float* Kblk) {
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
      int b = 0, slot = 0, pos = 0;
      bool tr = false, ok;
      if constexpr (KIND == FECX_TET10) {
        const uint32_t t = (stage_addr >> (8 * p)) & 0xFFu;
        ok = (t & 128u) != 0;
        tr = p == TET10_TR_JOB && (t & 64u) != 0;  // no other job is stored transposed (checked with the table)
        pos = (int)(t & 63u);
      } else {
        ok = PJ::job(lane, p, b, slot, tr);
        pos = pair_pos(NNPE, PJ::center(lane), b, slot);
      }
      if (ok) {
        float* dst = Kblk + pos * (NF * NF);
#pragma unroll
        for (int i = 0; i < NF; ++i)
#pragma unroll
          for (int k = 0; k < NF; ++k) dst[i * NF + k] = tr ? kreg[p][k * NF + i] : kreg[p][i * NF + k];
      }
    }
  };
  if constexpr (DO_D) {
    // diagonal of the element Jacobian (the AssembledDiagonal return type): lane a forms its own block K_aa only
    if (valid && lane < NNPE) {
      const int a = lane;
      float acc[NF * NF + M::KX] = {};
#pragma unroll 2
      for (int q = 0; q < NQ; ++q) {
        const auto& Q = work.qpb(grp).q[q];
        float C[M::NC];
        float va[M::NV], sa[M::NA];
        load_c(Q, C);
        node_vec(Q, C, a, va);
        M::template side<ND>(args.mp, C, qp_weight(Q), va, sa);
        if constexpr (DIAG0) M::template pair_diag<ND>(sa, va, acc);
        else M::template pair<ND>(sa, va, acc);
      }
      M::template finish<ND>(acc);
#pragma unroll
      for (int i = 0; i < NF; ++i) rel_out(a * NF + i, acc[i * NF + i]);
    }
  }
  if constexpr (DO_ASM) {
    // element action K_el v_el without atomics: stage the element's symmetric blocks in LDS (over the consumed
    // quadrature data), then lane a sums its row K_ab v_b over b in ascending order
    phase_sync<WORK>();
    if (valid && lane < PJ::LANES) stage_blocks(work.kb(grp));
    phase_sync<WORK>();
    if (valid && lane < NNPE) {
      float kv[NF] = {};
      uint32_t rows[3] = {0u, 0u, 0u};  // tet10: the lane's row of block addresses
      if constexpr (KIND == FECX_TET10) {
        const uint32_t* r = reinterpret_cast<const uint32_t*>(c_tet10_addr.rows[lane]);
        rows[0] = r[0]; rows[1] = r[1]; rows[2] = r[2];
      }
#pragma unroll
      for (int b = 0; b < NNPE; ++b) {
        int pos;
        bool ab;
        if constexpr (KIND == FECX_TET10) {
          const uint32_t t = (rows[b >> 2] >> (8 * (b & 3))) & 0xFFu;
          pos = (int)(t & 63u);
          ab = (t & 64u) != 0;
        } else {
          const PairRef pr = pair_ref(NNPE, lane, b);
          pos = pair_pos(NNPE, lane, b, pr.slot);
          ab = pr.ab;
        }
        const float* blk = work.kb(grp) + pos * (NF * NF);
#pragma unroll
        for (int i = 0; i < NF; ++i)
#pragma unroll
          for (int k = 0; k < NF; ++k) kv[i] = fmaf(ab ? blk[i * NF + k] : blk[k * NF + i], L.v[b][k], kv[i]);
      }
#pragma unroll
      for (int i = 0; i < NF; ++i) rel_out(lane * NF + i, kv[i]);
    }
  }
  if constexpr (DO_K) {
    // stage the blocks in LDS, over the consumed quadrature data
    phase_sync<WORK>();
    if (valid && lane < PJ::LANES) stage_blocks(work.kb(grp));
    phase_sync<WORK>();
  }
}

template <int KIND, int NF, int PHYS, int MODE, bool G64 = false>
__global__ __launch_bounds__(256) void element_kernel(ElementArgs args) {
  using T = ET<KIND>;
  using LDS = ElemLDS<KIND, NF, PHYS, MODE, G64>;
  constexpr int NNPE = T::NNPE, NQ = T::NQ, G = T::G, EPW = ELEM_THREADS / G, NDOF = NNPE * NF;
  constexpr bool DO_K = LDS::DO_K;
  constexpr int KW = LDS::KW;
  __shared__ LDS lds[EPW];
  __shared__ __attribute__((aligned(16))) WaveWork<LDS, EPW, NQ, G> work;
  const int grp = threadIdx.x / G;
  const int64_t e = args.ebeg + (int64_t)blockIdx.x * EPW + grp;  // element storage position
  const bool valid = e < args.ne;
  bool bad = false;
  const bool ld = valid && threadIdx.x % G < NNPE;
  const auto in = load_node_in<T::ND, NF>(args, ld ? (int64_t)args.conn[e * NNPE + threadIdx.x % G] : 0, ld,
                                          (MODE & (MODE_ACTION_MF | MODE_ACTION_ASM)) != 0);
  const int64_t sel = (PHYS == FECX_J2 && valid) ? (int64_t)args.eperm[e] : 0;
  element_eval<KIND, NF, PHYS, MODE>(args, e, sel, valid, in, lds, work,
                                     [&](int w, float v) { args.Rel[e * NDOF + w] = v; }, bad, threadIdx.x);
  if constexpr (DO_K) {
    // the whole wavefront copies its elements' contiguous storage range with consecutive lanes on consecutive words
    constexpr int EW = T::NPAIRS * NF * NF;   // words per element
    constexpr int EPWAVE = 64 / G;            // elements per wavefront
    const int wl = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int64_t e0 = args.ebeg + (int64_t)blockIdx.x * EPW + wv * EPWAVE;
    const int64_t nvalid = args.ne - e0 < EPWAVE ? args.ne - e0 : EPWAVE;
    static_assert(EW == KW && (EPWAVE * EW) % 4 == 0, "a wavefront's blocks start 16-byte aligned in LDS");
    if (nvalid > 0) {
      float* out = args.Kel + e0 * EW;
      const float* src = work.kb(wv * EPWAVE);
      const int total = (int)nvalid * EW;
      int done = 0;
      // nontemporal: the blocks are read once, by the reduction (measured -2.5 % on the reduction that follows)
      if (((e0 * EW) & 3) == 0) {  // 16-byte aligned in HBM too (always, unless a pipelined piece starts unaligned)
        done = total & ~3;
        constexpr int NIT = (EPWAVE * EW / 4 + 63) / 64;  // 16-byte pieces per lane for a full wavefront
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
          const int i = 4 * wl + 256 * it;
          if (i < done)
            __builtin_nontemporal_store(*reinterpret_cast<const f32x4*>(src + i), reinterpret_cast<f32x4*>(out + i));
        }
      }
      for (int idx = done + wl; idx < total; idx += 64) out[idx] = src[idx];
    }
  }
  if (bad) atomicOr(args.flag, 1);
}

template <int KIND, int NF, int PHYS, bool G64 = false>
static void launch_modes(int mode, const ElementArgs& a, hipStream_t s) {
  using T = ET<KIND>;
  constexpr int EPW = ELEM_THREADS / T::G;
  if (a.ne <= a.ebeg) return;
  const unsigned grid = (unsigned)((a.ne - a.ebeg + EPW - 1) / EPW);
  constexpr int B = ELEM_THREADS;
  switch (mode) {
    case MODE_R: element_kernel<KIND, NF, PHYS, MODE_R, G64><<<grid, B, 0, s>>>(a); break;
    case MODE_K: element_kernel<KIND, NF, PHYS, MODE_K, G64><<<grid, B, 0, s>>>(a); break;
    case MODE_R | MODE_K: element_kernel<KIND, NF, PHYS, MODE_R | MODE_K, G64><<<grid, B, 0, s>>>(a); break;
    case MODE_ACTION_MF: element_kernel<KIND, NF, PHYS, MODE_ACTION_MF, G64><<<grid, B, 0, s>>>(a); break;
    case MODE_ACTION_ASM: element_kernel<KIND, NF, PHYS, MODE_ACTION_ASM, G64><<<grid, B, 0, s>>>(a); break;
    case MODE_DIAG: element_kernel<KIND, NF, PHYS, MODE_DIAG, G64><<<grid, B, 0, s>>>(a); break;
    case MODE_QPGRAD:  // gradients do not depend on the physics: one instantiation per (KIND, NF)
      element_kernel<KIND, NF, NF == 1 ? FECX_POISSON : FECX_NEOHOOKEAN, MODE_QPGRAD, G64><<<grid, B, 0, s>>>(a);
      break;
    default: break;
  }
}

// ------------------------------------------------------------------ fused assembly of structured hex8 meshes
// On a structured hex8 mesh (StructuredMesh numbering, src/meshes/StructuredMesh.jl:85-131) the residual and the CSR
// Jacobian are assembled in one kernel, without the element-private storage of the two-phase path and its round trip
// through HBM (written once by the element kernel, gathered again by the reduction).
//
// Tiles: TX x TY node columns, each seen through the SX x SY = (TX + 1) x (TY + 1) element columns around it. The work
// is the list of (tile, node plane) units, tile-major; each work-group (one per slot the device can hold at once) takes
// one contiguous share of it and marches up through the element layers of each tile range [p0, p1) in its share,
// k = p0 - 1 .. p1 - 1. For each layer it evaluates the SX x SY elements of the tile (element_eval, the same code and
// bits as the element kernel) and leaves their blocks and rows in LDS. Node plane k touches layers k - 1 and k only,
// so after layer k:
//   - the rows of plane k are complete: each block is its partial sum over layer k - 1, kept in FP64 registers since
//     the previous layer, plus its blocks from layer k; they are written to the CSR values, each entry once;
//   - the layer-k blocks of plane k + 1 become the new partial sums (only neighbours one plane down or level with a
//     node have any).
// Reduction lanes: lane (dx, dy, node) owns the three blocks (dx, dy, dz = -1, 0, +1) of its node's row: per layer it
// writes the three of plane k and keeps two partial sums for plane k + 1. A wavefront's lanes share (dx, dy) but for
// one switch, so the element slots they read are the same up to the node's offset (no per-lane search); slots are
// WORD-strided by an odd count, so the lanes of one read hit distinct LDS banks. The lanes of (0, 0) also form the
// residual rows.
// Elements of the halo ring around a tile and of the layer below p0 are evaluated by both work-groups that need
// them. The FP64 sum of an entry runs over its element contributions of layer k - 1 in ascending element id, then over
// those of layer k: the contributions of the two-phase path, grouped by layer.
// Column slots come from a per-node mask of the 27 neighbours present in its CSR row (existing and free): neighbour
// offsets (dx, dy, dz) in lexicographic (dz, dy, dx) order are the row's ascending node order.
__device__ __forceinline__ int hex_local(int ox, int oy, int oz) {  // StructuredMesh.jl:116-123 local order
  return oz * 4 + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
}

template <int SX, int SY>
struct FusedTile {
  static constexpr int EPW = SX * SY, THREADS = 8 * EPW;  // element slots of a layer, 8 lanes each
  static constexpr int TX = SX - 1, TY = SY - 1, NODES = TX * TY;
  static constexpr int TASKS = 9 * NODES;  // reduction lanes (dx, dy, node), node fastest
  static constexpr int RES0 = THREADS - 64;  // residual lanes (node) in the last wavefront
  static_assert(TASKS <= RES0 && NODES <= 64, "one reduction task per lane, the residual lanes apart");
};
constexpr int FUSED_SX = 8, FUSED_SY = 8;

// LDS of the fused kernel: the element work area, reused between layers as the staging area of the plane's rows
// (node-major, each node's rows at NB * 27 words in their CSR layout), so that they leave as contiguous stores.
template <class LDS, int EPW, int KS, int NODES, int NB>
union FusedWork {
  ElemWork<LDS, EPW, 8, KS> work;
  double stage[NODES * 27 * NB];
};

template <int NF, int PHYS, int SX, int SY>
__global__ __launch_bounds__(8 * SX * SY) __attribute__((amdgpu_waves_per_eu(NF == 1 ? 4 : 2, 8)))
void fused_hex8_kernel(ElementArgs args, FusedArgs f) {
  using TL = FusedTile<SX, SY>;
  using LDS = ElemLDS<FECX_HEX8, NF, PHYS, MODE_R | MODE_K>;
  constexpr int EPW = TL::EPW, NB = NF * NF, RW = 27 * NB;
  constexpr int KS = LDS::KW | 1;  // odd: the slots read by one instruction fall in distinct banks
  constexpr int RS = 8 * NF + 1;
  __shared__ LDS lds[EPW];
  __shared__ __attribute__((aligned(16))) FusedWork<LDS, EPW, KS, TL::NODES, NB> fw;
  __shared__ float srow[EPW][RS];  // the layer's element rows (residual)
  __shared__ int2 out_rows[TL::NODES];  // per node of the plane being written: {first CSR block, words} or words 0
  __shared__ double pmbuf[NB][TL::TASKS];  // per reduction lane: the dz = -1 block of plane k + 1 (layer k only)
  auto& work = fw.work;
  const int tid = threadIdx.x, grp = tid / 8;
  const int sx = grp % SX, sy = grp / SX;  // this lane group's element slot
  const int nex = f.nx - 1, ney = f.ny - 1, nez = f.nz - 1;
  // this work-group's share [units b / G, units (b + 1) / G) of the units, in 32-bit arithmetic (a 64-bit division is
  // ~150 instructions): units = q G + rr, so units b / G = q b + rr b / G, where rr b < G^2 (units < 2^31 and
  // G < 2^16, both checked at launch)
  const uint32_t units = (uint32_t)f.tiles_x * (uint32_t)f.tiles_y * (uint32_t)f.nz;
  const uint32_t G = gridDim.x, uq = units / G, ur = units % G;
  uint32_t u = uq * blockIdx.x + ur * blockIdx.x / G;
  const uint32_t u_end = uq * (blockIdx.x + 1) + ur * (blockIdx.x + 1) / G;
  // reduction task of this lane: the blocks (dx, dy, *) of a node's row, or the node's residual row (the lanes of the
  // last wavefront, which also hand the plane's row ranges to the copy-out)
  const bool task = tid < TL::TASKS;
  const bool res_lane = tid >= TL::RES0 && tid < TL::RES0 + TL::NODES;
  const int node = task ? tid % TL::NODES : (res_lane ? tid - TL::RES0 : 0);
  // the lane groups' (dx, dy) in an order whose wavefronts straddle groups with overlapping candidate slots: a
  // wavefront runs the union of its lanes' candidates, 16 candidate rounds per layer in this order against 18 in
  // (dx, dy) order (swapping (-1, 0) and (1, 0))
  const int grp9 = tid / TL::NODES;
  const int dxdy = task ? (grp9 == 3 ? 5 : (grp9 == 5 ? 3 : grp9)) : 4;
  const int li = node % TL::TX, lj = node / TL::TX;
  const int dx = dxdy % 3 - 1, dy = dxdy / 3 - 1;
  const int dm = (dx + 1) + 3 * (dy + 1);  // neighbour index of (dx, dy, -1); + 9 and + 18 for dz = 0 and + 1
  bool bad = false;

  // the block K_ab of element slot Ke, a = the node at (ox, oy, oz), b = its neighbour (ox + dx, oy + dy, mz): its
  // words (as stored: K_ab, or K_ba when the result is false), then added to acc in FP64
  auto load_block = [&](const float* Ke, int a, int b, float (&v)[NB]) {
    const PairRef pr = pair_ref(8, a, b);
    const float* blk = Ke + pr.slot * NB;
#pragma unroll
    for (int c = 0; c < NB; ++c) v[c] = blk[c];
    return pr.ab;
  };
  auto add_block = [&](const float (&v)[NB], bool ab, double (&acc)[NB]) {
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
      for (int k = 0; k < NF; ++k) acc[i * NF + k] += (double)(ab ? v[i * NF + k] : v[k * NF + i]);
  };

  while (u < u_end) {
    const uint32_t tile = u / (uint32_t)f.nz;
    const int p0 = (int)(u % (uint32_t)f.nz);
    const int p1 = u_end - u < (uint32_t)(f.nz - p0) ? p0 + (int)(u_end - u) : f.nz;
    u += p1 - p0;
    const int x0 = (int)(tile % (uint32_t)f.tiles_x) * TL::TX, y0 = (int)(tile / (uint32_t)f.tiles_x) * TL::TY;
    const int ex = x0 - 1 + sx, ey = y0 - 1 + sy;
    const bool col_ok = ex >= 0 && ex < nex && ey >= 0 && ey < ney;
    const int gx = x0 + li, gy = y0 + lj;
    const bool node_ok = (task || res_lane) && gx < f.nx && gy < f.ny;
    // candidate slots around the node: bit 2 cx + cy for slot (li + cx, lj + cy), which holds the node at offsets
    // (1 - cx, 1 - cy) and the neighbour at (1 - cx + dx, 1 - cy + dy); both inside the element, element in the mesh
    unsigned cand = 0;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int cx = cc >> 1, cy = cc & 1, mx = 1 - cx + dx, my = 1 - cy + dy;
      const int cex = gx - 1 + cx, cey = gy - 1 + cy;
      if (mx >= 0 && mx <= 1 && my >= 0 && my <= 1 && cex >= 0 && cex < nex && cey >= 0 && cey < ney)
        cand |= 1u << cc;
    }
    double pz[NB], rp[NF];  // plane k + 1 partial sums: block dz = 0, residual (block dz = -1: pmbuf)
#pragma unroll
    for (int c = 0; c < NB; ++c) pz[c] = 0.0;
#pragma unroll
    for (int c = 0; c < NF; ++c) rp[c] = 0.0;
    // gather inputs of the lane's local node a = tid % 8 in its slot's element of layer k (node ids and element ids
    // follow from the structured numbering: no connectivity read)
    auto fetch = [&](int k) {
      const int a = tid % 8, oxa = (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0, oya = (a & 3) >= 2 ? 1 : 0;
      const bool ok = col_ok && k >= 0 && k < nez;
      const int64_t nd = (ex + oxa) + (int64_t)f.nx * ((ey + oya) + (int64_t)f.ny * (k + (a >> 2)));
      return load_node_in<3, NF>(args, ok ? nd : 0, ok, false);
    };
    NodeIn<3, NF> in = fetch(p0 - 1);
    // J2: the old state of the lane's quadrature point (lane tid % 8) in its slot's element of layer k, fetched with the
    // node inputs
    float so[7] = {};
    auto fetch_state = [&](int k) {
      if constexpr (PHYS == FECX_J2) {
        const bool ok = col_ok && k >= 0 && k < nez;
        const int64_t sidx = ok ? ((((int64_t)ex * ney + ey) * nez + k) * 8 + tid % 8) * 7 : 0;
#pragma unroll
        for (int st = 0; st < 7; ++st) so[st] = ok ? args.st_old[sidx + st] : 0.0f;
      }
    };
    fetch_state(p0 - 1);

    for (int k = p0 - 1; k < p1; ++k) {
      const bool layer_ok = k >= 0 && k < nez;
      const bool valid = layer_ok && col_ok;
      const int64_t eid = valid ? ((int64_t)ex * ney + ey) * nez + k : 0;
      const bool wr = node_ok && k >= p0;
      const int64_t n = gx + (int64_t)f.nx * (gy + (int64_t)f.ny * k);
      const bool wk = wr && f.vals;
      int2 info = make_int2(0, 0);
      if (wk) info = f.nodeinfo[n];  // read here, used after the element evaluation
      // the thread index through an empty asm statement: the element evaluation's per-lane constants (quadrature
      // point, shape functions) are then recomputed for every layer instead of held in registers across the loop
      int tid_k = tid;
      asm volatile("" : "+v"(tid_k));
      element_eval<FECX_HEX8, NF, PHYS, MODE_R | MODE_K, KS>(args, 0, eid, valid, in, lds, work,
                                                             [&](int w, float v) { srow[grp][w] = v; }, bad, tid_k,
                                                             PHYS == FECX_J2 ? so : nullptr);
      in = fetch(k + 1);  // the next layer's loads, in flight during the reduction
      fetch_state(k + 1);
      double a0[NB], a1[NB], pm[NB];  // plane k: blocks dz = 0, +1; plane k + 1: block dz = -1
      if (task) {
#pragma unroll
        for (int c = 0; c < NB; ++c) { a0[c] = pz[c]; a1[c] = 0.0; pm[c] = 0.0; pz[c] = 0.0; }
        if (layer_ok) {
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) {
            if ((cand >> cc) & 1u) {
              const int cx = cc >> 1, cy = cc & 1, ox = 1 - cx, oy = 1 - cy, mx = ox + dx, my = oy + dy;
              const float* Ke = work.K + ((li + cx) + SX * (lj + cy)) * KS;
              float v0[NB], v1[NB], v2[NB], v3[NB];  // the four blocks' loads in flight together
              const bool b0 = load_block(Ke, hex_local(ox, oy, 0), hex_local(mx, my, 0), v0);
              const bool b1 = load_block(Ke, hex_local(ox, oy, 0), hex_local(mx, my, 1), v1);
              const bool b2 = load_block(Ke, hex_local(ox, oy, 1), hex_local(mx, my, 0), v2);
              const bool b3 = load_block(Ke, hex_local(ox, oy, 1), hex_local(mx, my, 1), v3);
              add_block(v0, b0, a0);
              add_block(v1, b1, a1);
              add_block(v2, b2, pm);
              add_block(v3, b3, pz);
            }
          }
        }
      }
      if (res_lane) {
        {
          double r[NF];
#pragma unroll
          for (int c = 0; c < NF; ++c) { r[c] = rp[c]; rp[c] = 0.0; }
          if (layer_ok) {
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
              if ((cand >> cc) & 1u) {
                const int cx = cc >> 1, cy = cc & 1;
                const float* row = srow[(li + cx) + SX * (lj + cy)];
#pragma unroll
                for (int c = 0; c < NF; ++c) {
                  r[c] += (double)row[hex_local(1 - cx, 1 - cy, 0) * NF + c];
                  rp[c] += (double)row[hex_local(1 - cx, 1 - cy, 1) * NF + c];
                }
              }
            }
          }
          if (wr && f.R) {
#pragma unroll
            for (int c = 0; c < NF; ++c) f.R[n * NF + c] = r[c];
          }
        }
      }
      __syncthreads();  // every lane has read the element blocks: the work area becomes the staging area
      if (task || res_lane) {
        const uint32_t mask = (uint32_t)info.y;
        const int deg = __popc(mask);
        double* st = fw.stage + node * RW;
        auto put = [&](int d, const double (&acc)[NB]) {
          if ((mask >> d) & 1u) {
            double* dst = st + NF * __popc(mask & ((1u << d) - 1u));
#pragma unroll
            for (int i = 0; i < NF; ++i)
#pragma unroll
              for (int kk = 0; kk < NF; ++kk) dst[i * NF * deg + kk] = acc[i * NF + kk];
          }
        };
        if (task) {
          if (wk) {
            double am[NB];  // plane k, dz = -1: layer k - 1 only
#pragma unroll
            for (int c = 0; c < NB; ++c) am[c] = pmbuf[c][tid];
            put(dm, am);
            put(dm + 9, a0);
            put(dm + 18, a1);
          }
#pragma unroll
          for (int c = 0; c < NB; ++c) pmbuf[c][tid] = pm[c];
        } else {
          out_rows[node] = make_int2(info.x, wk ? NB * deg : 0);
        }
      }
      __syncthreads();
      // the next layer's inputs have arrived before the stores below are issued: the loads of the next element
      // evaluation then wait for no store
#pragma unroll
      for (int d = 0; d < 3; ++d) asm volatile("" ::"v"(in.hi[d]), "v"(in.lo[d]));
#pragma unroll
      for (int i = 0; i < NF; ++i) asm volatile("" ::"v"(in.u[i]));
      if constexpr (PHYS == FECX_J2) {
#pragma unroll
        for (int st = 0; st < 7; ++st) asm volatile("" ::"v"(so[st]));
      }
      // the plane's rows, node by node: each node's NB * deg words are contiguous in the CSR values
      {
        // wavefront w takes nodes w, w + W, ...; each node's rows in chunks of 64 words, every chunk's LDS read issued
        // before its stores
        constexpr int W = TL::THREADS / 64, NCH = (RW + 63) / 64, NPW = (TL::NODES + W - 1) / W;
        const int wave = tid / 64, wl = tid % 64;
#pragma unroll 1
        for (int t = 0; t < NPW; ++t) {
          const int nd = wave + W * t;
          const int2 o = nd < TL::NODES ? out_rows[nd] : make_int2(0, 0);
          const double* src = fw.stage + nd * RW;
          double v[NCH];
#pragma unroll
          for (int c = 0; c < NCH; ++c) v[c] = wl + 64 * c < o.y ? src[wl + 64 * c] : 0.0;
          double* dst = f.vals + (int64_t)NB * o.x;
#pragma unroll
          for (int c = 0; c < NCH; ++c)
            if (wl + 64 * c < o.y) __builtin_nontemporal_store(v[c], dst + wl + 64 * c);
        }
      }
      // no barrier here: element_eval's first barrier comes before its first write to the work area
    }
  }
  if (bad) atomicOr(args.flag, 1);
}

// The two scatter designs the fused kernel is measured against, on the same structured hex8 meshes and the same
// element code: every element's blocks and rows are added straight into the CSR values and the residual, which are
// zero-filled first. ATOMIC = 1: FP64 atomic adds (global_atomic_add_f64), all elements in one launch, each lane
// adding the 8 blocks of its element row (wavefronts hold 8 elements whose rows share nodes, so each entry still
// takes several atomics). ATOMIC = 0: element colouring -- 8 launches, one per colour (ex, ey, ez) mod 2, whose
// elements share no node, so plain read-add-write stores do not race. Summation order: atomics in arrival order
// (not deterministic), colouring in colour order.
template <int NF, int PHYS, int ATOMIC>
__global__ __launch_bounds__(ELEM_THREADS) __attribute__((amdgpu_waves_per_eu(2, 8)))
void scatter_hex8_kernel(ElementArgs args, FusedArgs f, int colour) {
  using LDS = ElemLDS<FECX_HEX8, NF, PHYS, MODE_R | MODE_K>;
  constexpr int EPW = ELEM_THREADS / 8, NB = NF * NF, KW = LDS::KW;
  __shared__ LDS lds[EPW];
  __shared__ __attribute__((aligned(16))) ElemWork<LDS, EPW, 8> work;
  __shared__ float srow[EPW][8 * NF];
  const int grp = threadIdx.x / 8, a = threadIdx.x % 8;
  const int nex = f.nx - 1, ney = f.ny - 1, nez = f.nz - 1;
  // element of this lane group: all elements in id order (atomics), or the colour's sub-grid; 32-bit index arithmetic
  // (fewer than 2^31 elements, checked at launch)
  const uint32_t idx = blockIdx.x * (uint32_t)EPW + (uint32_t)grp;
  int ex, ey, ez;
  bool valid;
  if (ATOMIC) {
    valid = idx < (uint32_t)nex * (uint32_t)ney * (uint32_t)nez;
    const uint32_t q = idx / (uint32_t)nez;
    ez = (int)(idx - q * (uint32_t)nez); ey = (int)(q % (uint32_t)ney); ex = (int)(q / (uint32_t)ney);
  } else {
    const int cx = colour & 1, cy = (colour >> 1) & 1, cz = (colour >> 2) & 1;
    const int mx = (nex - cx + 1) / 2, my = (ney - cy + 1) / 2, mz = (nez - cz + 1) / 2;
    valid = mx > 0 && my > 0 && mz > 0 && idx < (uint32_t)mx * (uint32_t)my * (uint32_t)mz;
    const uint32_t q = mz > 0 ? idx / (uint32_t)mz : 0;
    const int iz = mz > 0 ? (int)(idx - q * (uint32_t)mz) : 0, iy = my > 0 ? (int)(q % (uint32_t)my) : 0;
    const int ix = my > 0 ? (int)(q / (uint32_t)my) : 0;
    ex = 2 * ix + cx; ey = 2 * iy + cy; ez = 2 * iz + cz;
  }
  if (!valid) { ex = ey = ez = 0; }
  const int64_t e = valid ? (int64_t)f.epos[((int64_t)ex * ney + ey) * nez + ez] : 0;
  bool bad = false;
  const auto in = load_node_in<3, NF>(args, valid ? (int64_t)args.conn[e * 8 + threadIdx.x % 8] : 0, valid, false);
  const int64_t sel = valid ? ((int64_t)ex * ney + ey) * nez + ez : 0;
  element_eval<FECX_HEX8, NF, PHYS, MODE_R | MODE_K>(args, e, sel, valid, in, lds, work,
                                                       [&](int w, float v) { srow[grp][w] = v; }, bad, threadIdx.x);
  if (valid) {
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const int64_t na = (ex + off[a][0]) + (int64_t)f.nx * ((ey + off[a][1]) + (int64_t)f.ny * (ez + off[a][2]));
    if (f.R) {
#pragma unroll
      for (int i = 0; i < NF; ++i) {
        double* r = f.R + na * NF + i;
        const double v = (double)srow[grp][a * NF + i];
        if (ATOMIC) atomicAdd(r, v);
        else *r += v;
      }
    }
    const int2 info = f.nodeinfo[na];
    const uint32_t mask = (uint32_t)info.y;
    const int deg = __popc(mask);
    for (int b = 0; b < 8; ++b) {
      const int d = (off[b][2] - off[a][2] + 1) * 9 + (off[b][1] - off[a][1] + 1) * 3 + (off[b][0] - off[a][0] + 1);
      if (!((mask >> d) & 1u)) continue;
      const int slot = __popc(mask & ((1u << d) - 1u));
      const PairRef pr = pair_ref(8, a, b);
      const float* blk = work.K + grp * KW + pr.slot * NB;
      double* dst = f.vals + (int64_t)NB * info.x + NF * slot;
#pragma unroll
      for (int i = 0; i < NF; ++i)
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double* p = dst + (int64_t)i * NF * deg + k;
          const double v = (double)blk[pr.ab ? i * NF + k : k * NF + i];
          if (ATOMIC) atomicAdd(p, v);
          else *p += v;
        }
    }
  }
  if (bad) atomicOr(args.flag, 1);
}

// scatter 1: FP64 atomics, 2: element colouring; vals and R zero-filled by the caller
int launch_scatter_hex8(int nf, int scatter, const ElementArgs& a, const FusedArgs& f, hipStream_t s) {
  const int phys = a.mp.phys;
  const int64_t nex = f.nx - 1, ney = f.ny - 1, nez = f.nz - 1;
  constexpr int EPW = ELEM_THREADS / 8;
  auto go = [&](auto kern) {
    if (scatter == 1) {
      kern<<<(unsigned)((nex * ney * nez + EPW - 1) / EPW), ELEM_THREADS, 0, s>>>(a, f, 0);
    } else {
      for (int c = 0; c < 8; ++c) {
        const int64_t m = ((nex - (c & 1) + 1) / 2) * ((ney - ((c >> 1) & 1) + 1) / 2) * ((nez - ((c >> 2) & 1) + 1) / 2);
        if (m > 0) kern<<<(unsigned)((m + EPW - 1) / EPW), ELEM_THREADS, 0, s>>>(a, f, c);
      }
    }
  };
  if (scatter != 1 && scatter != 2) return -1;
  if (nex * ney * nez >= (int64_t(1) << 31)) return -1;  // the kernels' 32-bit element index
  if (nf == 1 && phys == FECX_POISSON) {
    if (scatter == 1) go(scatter_hex8_kernel<1, FECX_POISSON, 1>); else go(scatter_hex8_kernel<1, FECX_POISSON, 0>);
    return 0;
  }
  if (nf == 3 && phys == FECX_NEOHOOKEAN) {
    if (scatter == 1) go(scatter_hex8_kernel<3, FECX_NEOHOOKEAN, 1>); else go(scatter_hex8_kernel<3, FECX_NEOHOOKEAN, 0>);
    return 0;
  }
  return -1;
}

// One work-group per slot the device holds at once (the kernel's occupancy times the CU count), each with an equal
// share of the (tile, plane) units.
template <int NF, int PHYS>
static bool launch_fused(const ElementArgs& a, FusedArgs f, hipStream_t s) {
  using TL = FusedTile<FUSED_SX, FUSED_SY>;
  auto kern = fused_hex8_kernel<NF, PHYS, FUSED_SX, FUSED_SY>;
  static int slots = 0;  // per instantiation; the same on every gfx950 device
  if (slots == 0) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
        hipSuccess || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, TL::THREADS, 0) != hipSuccess)
      cus = per = 0;
    slots = (cus > 0 ? cus : 256) * (per > 0 ? per : 1);
  }
  f.tiles_x = (f.nx + TL::TX - 1) / TL::TX;
  f.tiles_y = (f.ny + TL::TY - 1) / TL::TY;
  const int64_t units = (int64_t)f.tiles_x * f.tiles_y * f.nz;
  const unsigned grid = (unsigned)(units < slots ? units : slots);
  if (units >= (int64_t(1) << 31) || grid >= (1u << 16)) return false;  // the kernel's 32-bit work split
  kern<<<grid, TL::THREADS, 0, s>>>(a, f);
  return true;
}

int launch_fused_hex8(int nf, const ElementArgs& a, const FusedArgs& f, hipStream_t s) {
  const int phys = a.mp.phys;
  bool ok = false;
  if (nf == 1 && phys == FECX_POISSON) ok = launch_fused<1, FECX_POISSON>(a, f, s);
  else if (nf == 3 && phys == FECX_LINEAR_ELASTIC) ok = launch_fused<3, FECX_LINEAR_ELASTIC>(a, f, s);
  else if (nf == 3 && phys == FECX_NEOHOOKEAN) ok = launch_fused<3, FECX_NEOHOOKEAN>(a, f, s);
  else if (nf == 3 && phys == FECX_J2) ok = launch_fused<3, FECX_J2>(a, f, s);
  return ok ? 0 : -1;
}

// tet10 in FP32 or with the FP64 element geometry (ElementArgs::fp64_geometry)
template <int NF, int PHYS>
static void launch_tet10(int mode, const ElementArgs& a, hipStream_t s) {
  if (a.fp64_geometry) launch_modes<FECX_TET10, NF, PHYS, true>(mode, a, s);
  else launch_modes<FECX_TET10, NF, PHYS, false>(mode, a, s);
}

int launch_element(int kind, int nf, int mode, const ElementArgs& a, hipStream_t s) {
  const int phys = a.mp.phys;
  if (nf == 1 && phys == FECX_POISSON) {
    if (kind == FECX_QUAD4) { launch_modes<FECX_QUAD4, 1, FECX_POISSON>(mode, a, s); return 0; }
    if (kind == FECX_HEX8) { launch_modes<FECX_HEX8, 1, FECX_POISSON>(mode, a, s); return 0; }
    if (kind == FECX_TET10) { launch_tet10<1, FECX_POISSON>(mode, a, s); return 0; }
    return -1;
  }
  if (nf == 2 && kind == FECX_QUAD4 && phys == FECX_LINEAR_ELASTIC) {  // plane strain
    launch_modes<FECX_QUAD4, 2, FECX_LINEAR_ELASTIC>(mode, a, s);
    return 0;
  }
  if (nf == 3 && (kind == FECX_HEX8 || kind == FECX_TET10)) {
    if (kind == FECX_HEX8) {
      if (phys == FECX_LINEAR_ELASTIC) { launch_modes<FECX_HEX8, 3, FECX_LINEAR_ELASTIC>(mode, a, s); return 0; }
      if (phys == FECX_NEOHOOKEAN) { launch_modes<FECX_HEX8, 3, FECX_NEOHOOKEAN>(mode, a, s); return 0; }
      if (phys == FECX_J2) { launch_modes<FECX_HEX8, 3, FECX_J2>(mode, a, s); return 0; }
      if (phys == FECX_J2_SURROGATE) { launch_modes<FECX_HEX8, 3, FECX_J2_SURROGATE>(mode, a, s); return 0; }
      if (phys == FECX_NH_SURROGATE) { launch_modes<FECX_HEX8, 3, FECX_NH_SURROGATE>(mode, a, s); return 0; }
    } else {
      if (phys == FECX_LINEAR_ELASTIC) { launch_tet10<3, FECX_LINEAR_ELASTIC>(mode, a, s); return 0; }
      if (phys == FECX_NEOHOOKEAN) { launch_tet10<3, FECX_NEOHOOKEAN>(mode, a, s); return 0; }
      if (phys == FECX_J2) { launch_tet10<3, FECX_J2>(mode, a, s); return 0; }
      if (phys == FECX_J2_SURROGATE) { launch_tet10<3, FECX_J2_SURROGATE>(mode, a, s); return 0; }
      if (phys == FECX_NH_SURROGATE) { launch_tet10<3, FECX_NH_SURROGATE>(mode, a, s); return 0; }
    }
  }
  return -1;
}

}  // namespace fecx
