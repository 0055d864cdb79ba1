// Reduction kernels (phase 2 of assembly): FP64 sums of the FP32 element-private outputs.
//
// Matrix: the reference reduces its element-private COO storage with SparseArrays.sparse! over the sorted
// permutation on every stiffness() call (src/assemblers/SparsityPatterns.jl:301-329). Here a group of LPR lanes
// owns one node: lane s owns column slot s of the node's NF CSR rows, walks the node's incident elements in
// ascending element order (the order sparse! sums duplicates in), adds the NF x NF blocks that land in its slot
// in FP64, and writes them with plain stores -- no atomics, no zero fill, every CSR value written exactly once,
// adjacent lanes writing adjacent columns.
// Vector: the same groups sum the node's incident element rows into the residual field (the atomic scatter of
// src/assemblers/Assemblers.jl:72-87 restated as an owner-computes gather).
// These kernels share this contract (launch_buckets routes the light rows, then the heavy ones):
//   reduce_scalar_hex8_kernel   scalar hex8 light rows: lanes add their words into LDS slot accumulators;
//   reduce_scatter_kernel       3-field light rows of up to 32 blocks, and the tet10 heavy rows (scalar and 3-field)
//                               of up to 72 blocks: the same scatter, 32 or 64 lanes per row;
//   reduce_merged_kernel        3-field tet10 meshes whose rows have at most 65 blocks: both buckets of
//                               reduce_scatter_kernel in one launch, their work-groups interleaved along the Morton curve;
//   reduce_rows_kernel          slot-owner lanes for everything else: light scalar tet10/quad4 rows, 3-field light
//                               rows of more than 32 blocks, heavy hex8/quad4 rows and tet10 rows of more than 72;
//   reduce_vector_kernel / reduce_vector_group_kernel   calls without CSR values (one thread, or 16 lanes, per row).
// All sum a slot's (or a row's) contributions in ascending incidence order, so every form gives the same bits.
#include <algorithm>

#include "fecx_internal.h"

namespace fecx {

constexpr int RED_THREADS = 256;
constexpr int CH = 8;            // incidences staged per pass
constexpr int RED_NF1_LPR = 8;     // lanes per scalar row of at most 32 blocks
constexpr int RED_NF3_LPR = 32;    // lanes per vector row of at most 32 blocks
constexpr int RED_HEAVY_LPR = 128;  // lanes per row of 65..256 blocks (one slot per lane and sweep, two sweeps past 128)
constexpr int RED_LB = 4;          // vector rows: incidences whose block loads are in flight together
constexpr int RED_LB1 = 8;         // the same for rows with one staged word per lane and incidence
constexpr int RED_HEAVY_CH = 24;   // incidences staged per pass on 128-lane rows (tet10 interior vertices: 24)

// Position of byte == s in the packed slot words (at most one match: a node appears once per element), or -1.
template <int VW>
__device__ __forceinline__ int match_slot(const uint32_t* w, int s) {
  const uint32_t rep = 0x01010101u * (uint32_t)s;
#pragma unroll
  for (int q = 0; q < VW; ++q) {
    const uint32_t x = w[q] ^ rep;
    const uint32_t m = (x - 0x01010101u) & ~x & 0x80808080u;  // lowest flagged byte is the first zero byte
    if (m) return 4 * q + (__builtin_ctz(m) >> 3);
  }
  return -1;
}

// A work-group owns 256/LPR consecutive nodes (after the XCD remap). Per pass, a node's LPR lanes
//   1. stage up to CHK incidences (element storage position, local node a, column slots, residual row) in LDS:
//      8, or 24 on the 128-lane rows (a tet10 interior vertex has 24 incident elements: one pass instead of three),
//   2. copy the NNPE blocks K_ab (b = 0..NNPE-1) of each staged incidence from the symmetric element storage into
//      LDS; a per-work-group table gives, for (a, word j of the staged row), the word's offset inside the element's
//      storage with the transposition of blocks stored as K_ba folded in,
//   3. sum in FP64, lane s owning column slot s, the staged blocks that land in slot s, in incidence order.
// The residual sum runs over the same staged rows in the same order.
// LDS of reduce_rows_kernel<NF, NNPE, LPR, SW, CHK> (the __shared__ arrays below, plus the pass-count word).
template <int NF, int NNPE, int LPR, int CHK>
constexpr int reduce_rows_lds() {
  constexpr int NPW = RED_THREADS / LPR, BW = NNPE * NF * NF, VW = (NNPE + 3) / 4;
  return NNPE * BW * 2 + NPW * CHK * 4 * (1 + VW + NF + BW) + 4;
}
// Vector rows ask for 7 waves per SIMD (at most 72 VGPRs) where LDS leaves room for 7 work-groups per CU (under
// 22.5 KB each in 2 KB granules): the 32-lane hex8 kernel then holds 70 VGPRs instead of 77, one occupancy step up,
// and its reduction measured 1.043 against 1.084 ms on neo-Hookean hex8 96^3. The scalar kernels already fit 8 waves
// (the tet10 one would spill under the cap); the tet10 32-lane kernel is held to 6 by its LDS.
template <int NF, int NNPE, int LPR, int SW, int CHK>
constexpr int reduce_rows_waves() {
  return (NF == 3 && SW == 1 && reduce_rows_lds<NF, NNPE, LPR, CHK>() <= 22528) ? 7 : 1;
}
template <int NF, int NNPE, int LPR, int SW, int CHK = CH>
__global__ __launch_bounds__(RED_THREADS)
__attribute__((amdgpu_waves_per_eu(reduce_rows_waves<NF, NNPE, LPR, SW, CHK>(), 8)))
void reduce_rows_kernel(ReduceArgs a) {
  static_assert(CHK <= LPR && CHK % 8 == 0, "a row's lanes stage its incidences, 8 per batch");
  constexpr int NPW = RED_THREADS / LPR;  // nodes per work-group
  constexpr int NPAIRS = NNPE * (NNPE + 1) / 2;
  constexpr int EW = NPAIRS * NF * NF;    // stored words per element
  constexpr int VW = (NNPE + 3) / 4;      // vslot words per incidence
  constexpr int BW = NNPE * NF * NF;      // staged words per incidence
  __shared__ int16_t s_off[NNPE][BW];
  __shared__ int32_t s_v[NPW][CHK];
  __shared__ uint32_t s_w[NPW][CHK][VW];
  __shared__ float s_r[NPW][CHK][NF];
  __shared__ float s_blk[NPW][CHK][BW];
  __shared__ int s_maxpass;
  for (int i = threadIdx.x; i < NNPE * BW; i += RED_THREADS) (&s_off[0][0])[i] = a.kel_off[i];
  const int64_t chunk = xcd_chunk(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
  const int64_t pos = a.p0 + chunk * NPW + grp;
  const bool valid = pos < a.p1;
  // one 16-byte row descriptor: {node, first incidence, first block of the row, incidences | degree << 16}
  const int4 ri = valid ? a.rowinfo[pos] : make_int4(0, 0, 0, 0);
  const int64_t n = ri.x;
  const int32_t c0 = ri.y, ninc = ri.w & 0xFFFF;
  int maxpass = 1;  // light rows (at most CHK incidences) take one pass
  if (a.onepass) {
    __syncthreads();  // s_off
  } else {
    const int npass = (ninc + CHK - 1) / CHK;
    if (threadIdx.x == 0) s_maxpass = 0;
    __syncthreads();
    if (lane == 0) atomicMax(&s_maxpass, npass);
    __syncthreads();
    maxpass = s_maxpass;
  }
  // a row of a node with Dirichlet dofs has no CSR blocks (degree 0); a free node always couples with itself
  const int deg = a.vals ? (ri.w >> 16) : 0;
  const int32_t r = deg > 0 ? 0 : -1;
  const int64_t rb = ri.z;
  double acc[SW][NF * NF];
#pragma unroll
  for (int t = 0; t < SW; ++t)
#pragma unroll
    for (int c = 0; c < NF * NF; ++c) acc[t][c] = 0.0;
  double racc = 0.0;
  for (int pass = 0; pass < maxpass; ++pass) {
    const int base = pass * CHK;
    const int cnt = min(CHK, ninc - base);  // may be <= 0 for this node
    // padded light rows (scalar kernels only): the 8 slots at 8 pos are loaded without waiting for the row descriptor
    // (unused slots hold a valid incidence and are never summed: every later loop stops at cnt). The vector-row
    // kernels keep the plain form, which holds their register count below the next occupancy step.
    bool stage;
    int64_t c;
    if constexpr (NF == 1) {
      stage = (CHK == CH && a.padded) ? (valid && lane < CH) : (lane < cnt);
      c = (CHK == CH && a.padded) ? pos * CH + lane : (int64_t)c0 + base + lane;
    } else {
      stage = lane < cnt;
      c = c0 + base + lane;
    }
    if (stage) {
      const int32_t v = a.vent[c];
      s_v[grp][lane] = v;
      const uint32_t* sl = reinterpret_cast<const uint32_t*>(a.vslot) + (int64_t)c * VW;
#pragma unroll
      for (int q = 0; q < VW; ++q) s_w[grp][lane][q] = sl[q];
      if (a.R) {
#pragma unroll
        for (int i = 0; i < NF; ++i) s_r[grp][lane][i] = a.Rel[(int64_t)v * NF + i];
      }
    }
    __syncthreads();
    if (r >= 0) {
      if constexpr (NF == 1 && LPR > BW) {
        // scalar rows on more lanes than words per incidence: one load per lane and incidence
        for (int t = 0; t < cnt; ++t) {
          const uint32_t v = (uint32_t)s_v[grp][t];
          const uint32_t ebase = (v / NNPE) * (uint32_t)EW;  // element storage is limited to 2^31 words
          const int la = (int)(v % NNPE);
#pragma unroll
          for (int j0 = 0; j0 < BW; j0 += LPR) {
            const int j = j0 + lane;
            if (j < BW) s_blk[grp][t][j] = a.Kel[ebase + (uint32_t)s_off[la][j]];
          }
        }
      } else {
        // vector rows: incidences in batches of LB -- all word indices of a batch first, then its loads back to back
        // (their latencies overlap), then the LDS writes. Slots past the row's incidences re-read the first
        // incidence's base word and are never written. With three words per lane and incidence (32 lanes), LB = 4
        // keeps the kernel near 80 VGPRs (6 waves per SIMD); 8 in flight needs ~106 and measured slower. With one word
        // per lane (scalar rows, 128-lane rows) 8 incidences cost 16 VGPRs and are all in flight.
        constexpr int JN = (BW + LPR - 1) / LPR;
        constexpr int LB = (NF == 1 || JN == 1) ? RED_LB1 : RED_LB;
#pragma unroll
        for (int t0 = 0; t0 < CHK; t0 += LB) {
          if (t0 >= cnt) break;  // no incidence left in this batch (s_v[grp][0] is set whenever t0 < cnt)
          uint32_t idx[LB][JN];
#pragma unroll
          for (int tt = 0; tt < LB; ++tt) {
            const int t = t0 + tt;
            const uint32_t v = (uint32_t)s_v[grp][t < cnt ? t : 0];
            const uint32_t ebase = (v / NNPE) * (uint32_t)EW;
            const int la = (int)(v % NNPE);
#pragma unroll
            for (int jj = 0; jj < JN; ++jj) {
              const int j = jj * LPR + lane;
              idx[tt][jj] = (t < cnt && j < BW) ? ebase + (uint32_t)s_off[la][j] : ebase;
            }
          }
          float tmp[LB][JN];
#pragma unroll
          for (int tt = 0; tt < LB; ++tt)
#pragma unroll
            for (int jj = 0; jj < JN; ++jj) tmp[tt][jj] = a.Kel[idx[tt][jj]];
#pragma unroll
          for (int tt = 0; tt < LB; ++tt) {
            const int t = t0 + tt;
            if (t < cnt) {
#pragma unroll
              for (int jj = 0; jj < JN; ++jj) {
                const int j = jj * LPR + lane;
                if (j < BW) s_blk[grp][t][j] = tmp[tt][jj];
              }
            }
          }
        }
      }
    }
    __syncthreads();
    if (a.R && lane < NF) {
      for (int t = 0; t < cnt; ++t) racc += (double)s_r[grp][t][lane];
    }
    if (r >= 0) {
#pragma unroll
      for (int sw = 0; sw < SW; ++sw) {
        const int s = lane + sw * LPR;
        if (s < deg) {
          for (int t = 0; t < cnt; ++t) {
            const int b = match_slot<VW>(s_w[grp][t], s);
            if (b >= 0) {
              const float* blk = &s_blk[grp][t][b * NF * NF];
#pragma unroll
              for (int c = 0; c < NF * NF; ++c) acc[sw][c] += (double)blk[c];
            }
          }
        }
      }
    }
    if (pass + 1 < maxpass) __syncthreads();  // the next pass restages s_v, s_w, s_r and s_blk
  }
  if (valid && a.R && lane < NF) a.R[n * NF + lane] = racc;
  if (r >= 0) {
#pragma unroll
    for (int sw = 0; sw < SW; ++sw) {
      const int s = lane + sw * LPR;
      if (s < deg) {
        const int64_t base = (int64_t)NF * NF * rb + (int64_t)NF * s;
#pragma unroll
        for (int i = 0; i < NF; ++i)
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            double* dst = &a.vals[base + (int64_t)i * NF * deg + k];
            if constexpr (NF >= 3) __builtin_nontemporal_store(acc[sw][i * NF + k], dst);  // written once, never read here
            else *dst = acc[sw][i * NF + k];
          }
      }
    }
  }
}

// Scalar hex8 light rows (the Poisson matrix): the gather leaves word b of every staged incidence in lane b's
// registers, so each lane adds its words straight into FP64 accumulators of the row's column slots in LDS, one
// incidence after the other in ascending order. Within one incidence the lanes hit distinct slots (a node appears once
// per element), so the plain read-add-write steps never collide, and each slot sums its blocks in the order of the
// slot-owner kernel above, with the same bits, but without that kernel's scan of every incidence's slot bytes for
// every slot (814 VALU instructions per wavefront there, most of its time on the Poisson 128^3 matrix).
template <int SMAX>
__global__ __launch_bounds__(RED_THREADS) void reduce_scalar_hex8_kernel(ReduceArgs a) {
  constexpr int NNPE = 8, LPR = 8, NPW = RED_THREADS / LPR, EW = NNPE * (NNPE + 1) / 2;
  static_assert(CH == LPR, "one staged incidence per lane");
  __shared__ int16_t s_off[NNPE][NNPE];
  __shared__ int32_t s_v[NPW][CH];
  __shared__ uint32_t s_w[NPW][CH][2];
  __shared__ float s_r[NPW][CH];
  __shared__ double s_acc[NPW][SMAX + 1];  // + 1: the rows of one wavefront start on different banks
  for (int i = threadIdx.x; i < NNPE * NNPE; i += RED_THREADS) (&s_off[0][0])[i] = a.kel_off[i];
  const int64_t chunk = xcd_chunk(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
  const int64_t pos = a.p0 + chunk * NPW + grp;
  const bool valid = pos < a.p1;
  const int4 ri = valid ? a.rowinfo[pos] : make_int4(0, 0, 0, 0);
  const int64_t n = ri.x;
  const int32_t c0 = ri.y, ninc = ri.w & 0xFFFF;
  const int deg = ri.w >> 16;  // 0 for a node with Dirichlet dofs
  const int cnt = min(CH, ninc);
#pragma unroll
  for (int k = lane; k < SMAX; k += LPR) s_acc[grp][k] = 0.0;
  // stage the incidences (padded rows: the 8 slots at 8 pos, loaded without waiting for the descriptor)
  const bool stage = a.padded ? (valid && lane < CH) : (lane < cnt);
  const int64_t c = a.padded ? pos * CH + lane : (int64_t)c0 + lane;
  if (stage) {
    const int32_t v = a.vent[c];
    s_v[grp][lane] = v;
    const uint32_t* sl = reinterpret_cast<const uint32_t*>(a.vslot) + c * 2;
    s_w[grp][lane][0] = sl[0];
    s_w[grp][lane][1] = sl[1];
    if (a.R) s_r[grp][lane] = a.Rel[v];
  }
  __syncthreads();
  if (deg > 0 && cnt > 0) {
    float w[CH];  // word `lane` (block K_a,lane) of every staged incidence, all loads in flight
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const uint32_t v = (uint32_t)s_v[grp][t < cnt ? t : 0];
      w[t] = a.Kel[(v / NNPE) * (uint32_t)EW + (uint32_t)s_off[v % NNPE][lane]];
    }
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      if (t < cnt) {
        const uint32_t sb = (s_w[grp][t][lane >> 2] >> (8 * (lane & 3))) & 0xFFu;  // this column's slot in the row
        if (sb != 0xFFu) s_acc[grp][sb] += (double)w[t];
      }
    }
  }
  __syncthreads();
  if (a.R && lane == 0) {
    double racc = 0.0;
    for (int t = 0; t < cnt; ++t) racc += (double)s_r[grp][t];
    if (valid) a.R[n] = racc;
  }
  for (int k = lane; k < deg; k += LPR) a.vals[(int64_t)ri.z + k] = s_acc[grp][k];
}

// Vector rows as a scatter, LPR lanes per row within one wavefront (light rows: 32 lanes, at most 8 incidences and 32
// blocks; tet10 vertex rows: 64 lanes, 24 incidences per pass, 65 blocks). Lane l holds words l, l + LPR, ... of the
// row's staged K_a* (block b = word / NF^2, entry c = word % NF^2) for each incidence, loaded a batch of incidences
// at a time, and adds each into the FP64 accumulator of (slot of b, c) in LDS, incidence after incidence in ascending
// order. Within one incidence the blocks land in distinct slots, and the LDS operations of one wavefront complete in
// order, so every accumulator sums its words in the slot-owner kernel's order, with the same bits. The slot-owner
// form tests every staged block's slot byte for every slot (240 tests per lane on a tet10 vertex row) and runs a
// block's 9 FP64 adds on every lane whenever one lane matched: about 1,300 instructions per wavefront on the vertex
// rows, which bound its time. Rows of more than SMAX blocks keep the slot-owner kernel.
template <int NF, int NNPE, int LPR, int CHK, int SMAX>
struct ScatterLDS {
  static constexpr int NPW = RED_THREADS / LPR, NB = NF * NF, VW = (NNPE + 3) / 4, BW = NNPE * NB;
  // block (a, b) of the staged rows: first word of its block in the element storage, bit 15 set when the storage holds
  // K_ba (the words then transposed); 2 * NNPE^2 bytes instead of 2 * NNPE * BW for a table of every word
  uint16_t s_blk[NNPE][NNPE];
  int32_t s_v[NPW][CHK];
  uint32_t s_w[NPW][CHK][VW];
  // staged residual rows; 64-lane rows (one per wavefront, so a uniform incidence index) read them from the staging
  // lanes' registers instead
  static constexpr bool STAGE_R = LPR < 64;
  float s_r[STAGE_R ? NPW : 1][STAGE_R ? CHK : 1][NF];
  double s_acc[NPW][SMAX * NB];
};
// A row's lanes lie in one wavefront and share LDS with no other row: between its staging and its sums, ordering
// comes from the program order of the wavefront's LDS operations, which complete in order, kept by a wavefront-scope
// fence, so the rows of a work-group never wait for each other.
__device__ __forceinline__ void row_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// The work-group's rows are the processing positions first + (thread / LPR), those below `end`.
template <int NF, int NNPE, int LPR, int CHK, int SMAX>
__device__ __forceinline__ void scatter_rows(const ReduceArgs& a, int64_t first, int64_t end, bool onepass,
                                             ScatterLDS<NF, NNPE, LPR, CHK, SMAX>& L) {
  static_assert(LPR <= 64 && 64 % LPR == 0, "a row's lanes lie in one wavefront");
  constexpr int NPW = RED_THREADS / LPR;
  constexpr int NB = NF * NF;                      // words per block
  constexpr int EW = NNPE * (NNPE + 1) / 2 * NB;   // stored words per element
  constexpr int VW = (NNPE + 3) / 4;               // vslot words per incidence
  constexpr int BW = NNPE * NB;                    // staged words per incidence
  constexpr int JN = (BW + LPR - 1) / LPR;         // words per lane and incidence
  // incidences whose loads are in flight together: 2 on tet10 light rows (3 words per lane; neo-Hookean tet10 80^3 x 6
  // reduction 5.72 against 5.85 ms with 4), 4 on hex8 ones (neo-Hookean hex8 96^3: 0.983 against 1.001 ms with 2)
  constexpr int LB = JN <= 2 ? RED_LB1 : (NNPE == 10 ? 2 : RED_LB);
  static_assert(CHK <= LPR && CHK % LB == 0, "a row's lanes stage its incidences, LB per batch");
  auto& s_blk = L.s_blk;
  auto& s_v = L.s_v;
  auto& s_w = L.s_w;
  auto& s_r = L.s_r;
  constexpr bool STAGE_R = ScatterLDS<NF, NNPE, LPR, CHK, SMAX>::STAGE_R;
  auto& s_acc = L.s_acc;
  for (int i = threadIdx.x; i < NNPE * NNPE; i += RED_THREADS) {  // from the word table: words 0 and 1 of each block
    const int16_t* o = a.kel_off + (i / NNPE) * BW + (i % NNPE) * NB;
    const bool tr = NF > 1 && o[1] != o[0] + 1;
    (&s_blk[0][0])[i] = (uint16_t)(o[0] | (tr ? 0x8000 : 0));
  }
  const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
  const int64_t pos = first + grp;
  const bool valid = pos < end;
  const int4 ri = valid ? a.rowinfo[pos] : make_int4(0, 0, 0, 0);
  const int64_t n = ri.x;
  const int32_t c0 = ri.y, ninc = ri.w & 0xFFFF;
  const int deg = a.vals ? (ri.w >> 16) : 0;  // 0: no CSR blocks (a node with Dirichlet dofs)
  for (int k = lane; k < deg * NB; k += LPR) s_acc[grp][k] = 0.0;
  // passes of this row (light rows, at most CHK incidences, take one); the two rows of a 32-lane wavefront may differ
  const int maxpass = onepass ? 1 : (ninc + CHK - 1) / CHK;
  __syncthreads();  // s_blk, the work-group's only barrier
  // this lane's words: block and entry of word jj * LPR + lane, and the entry's place in the transposed block
  int wb[JN], wc[JN], wt[JN];
#pragma unroll
  for (int jj = 0; jj < JN; ++jj) {
    const int j = jj * LPR + lane;
    wb[jj] = j < BW ? j / NB : 0;
    wc[jj] = j % NB;
    wt[jj] = (wc[jj] % NF) * NF + wc[jj] / NF;
  }
  double racc = 0.0;
  for (int pass = 0; pass < maxpass; ++pass) {
    const int base = pass * CHK;
    const int cnt = min(CHK, ninc - base);  // may be <= 0 for this node
    float rr[NF] = {};  // 64-lane rows: this lane's incidence's residual row
    if (lane < cnt) {
      const int64_t c = (int64_t)c0 + base + lane;
      const int32_t v = a.vent[c];
      s_v[grp][lane] = v;
      const uint32_t* sl = reinterpret_cast<const uint32_t*>(a.vslot) + c * VW;
#pragma unroll
      for (int q = 0; q < VW; ++q) s_w[grp][lane][q] = sl[q];
      if (a.R) {
#pragma unroll
        for (int i = 0; i < NF; ++i) {
          const float r = a.Rel[(int64_t)v * NF + i];
          if constexpr (STAGE_R) s_r[grp][lane][i] = r;
          else rr[i] = r;
        }
      }
    }
    row_sync();  // s_v, s_w and s_r staged by the row's lanes
    if constexpr (STAGE_R) {
      if (a.R && lane < NF) {
        for (int t = 0; t < cnt; ++t) racc += (double)s_r[grp][t][lane];
      }
    } else if (a.R) {
      for (int t = 0; t < cnt; ++t) {  // lane i < NF adds entry i of incidence t's row, held by lane t
        float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rr[0]), t));
#pragma unroll
        for (int i = 1; i < NF; ++i) {
          const float xi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rr[i]), t));
          if (lane == i) x = xi;
        }
        racc += (double)x;
      }
    }
    if (deg > 0) {
#pragma unroll
      for (int t0 = 0; t0 < CHK; t0 += LB) {
        if (t0 >= cnt) break;
        float w[LB][JN];  // all loads of the batch in flight together
#pragma unroll
        for (int tt = 0; tt < LB; ++tt) {
          const int t = t0 + tt;
          const uint32_t v = (uint32_t)s_v[grp][t < cnt ? t : 0];
          const uint32_t ebase = (v / NNPE) * (uint32_t)EW;  // element storage is limited to 2^31 words
          const int la = (int)(v % NNPE);
#pragma unroll
          for (int jj = 0; jj < JN; ++jj) {
            const uint32_t e = s_blk[la][wb[jj]];
            const uint32_t off = (e & 0x7FFFu) + (uint32_t)((e & 0x8000u) ? wt[jj] : wc[jj]);
            w[tt][jj] = a.Kel[ebase + (jj * LPR + lane < BW ? off : 0u)];
          }
        }
#pragma unroll
        for (int tt = 0; tt < LB; ++tt) {
          const int t = t0 + tt;
          if (t < cnt) {
#pragma unroll
            for (int jj = 0; jj < JN; ++jj) {
              if (jj * LPR + lane < BW) {
                const int b = wb[jj];
                const uint32_t sb = (s_w[grp][t][b >> 2] >> (8 * (b & 3))) & 0xFFu;  // the block's slot in the row
                if (sb != 0xFFu) s_acc[grp][sb * NB + wc[jj]] += (double)w[tt][jj];
              }
            }
          }
        }
      }
    }
    row_sync();  // the next pass restages s_v, s_w and s_r
  }
  if (valid && a.R && lane < NF) a.R[n * NF + lane] = racc;
  // one wavefront wrote this row's accumulators, in order: no barrier before reading them back
  const int64_t rb = ri.z;
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    double* dst = a.vals + (int64_t)NB * rb + (int64_t)i * NF * deg;
    for (int m = lane; m < NF * deg; m += LPR) {
      const double x = s_acc[grp][(m / NF) * NB + i * NF + m % NF];
      if constexpr (NF >= 3) __builtin_nontemporal_store(x, dst + m);  // written once, never read here
      else dst[m] = x;
    }
  }
}

template <int NF, int NNPE, int LPR, int CHK, int SMAX>
__global__ __launch_bounds__(RED_THREADS) void reduce_scatter_kernel(ReduceArgs a) {
  __shared__ ScatterLDS<NF, NNPE, LPR, CHK, SMAX> L;
  constexpr int NPW = RED_THREADS / LPR;
  scatter_rows<NF, NNPE, LPR, CHK, SMAX>(a, a.p0 + xcd_chunk(blockIdx.x, gridDim.x) * NPW, a.p1, a.onepass != 0, L);
}

constexpr int HEAVY_SCATTER_SMAX = 72;  // structured tet10 vertex rows: 65 blocks
constexpr int MERGED_HEAVY_SMAX = 65;   // the same rows in reduce_merged_kernel, whose LDS then fits 8 work-groups per CU

// tet10 rows of 3 fields, light (edge) and heavy (vertex) rows in one launch: work-group w takes merged chunk w (after
// the XCD remap), RED_MERGE_LIGHT edge rows on 32 lanes each or RED_MERGE_HEAVY vertex rows on 64 lanes each (the two
// bucket kernels' forms, LDS shared between them), the chunks ordered by their first rows' Morton ranks. A bucket's
// own launch reads the blocks its rows share with the other bucket's rows (24 of an element's 55) a second time,
// long after the first; here the vertex and edge rows of the same elements run at about the same time on the same
// XCD. Every row's sums are those of the bucket kernels, so the bits are too.
__global__ __launch_bounds__(RED_THREADS) __attribute__((amdgpu_waves_per_eu(8, 8))) void reduce_merged_kernel(
    ReduceArgs a) {
  static_assert(RED_THREADS / 32 == RED_MERGE_LIGHT && RED_THREADS / 64 == RED_MERGE_HEAVY, "chunk sizes");
  __shared__ union {
    ScatterLDS<3, 10, 32, CH, 32> lo;
    ScatterLDS<3, 10, 64, RED_HEAVY_CH, MERGED_HEAVY_SMAX> hi;
  } L;
  static_assert(sizeof(L) <= 163840 / 8, "8 work-groups per CU");
  const int64_t c = xcd_chunk(blockIdx.x, gridDim.x);
  if (c >= a.nmerged) return;  // the grid is rounded up to a multiple of 8
  const int2 ch = a.merged[c];
  if (ch.y == 0) scatter_rows<3, 10, 32, CH, 32>(a, ch.x, a.nlight, false, L.lo);
  else scatter_rows<3, 10, 64, RED_HEAVY_CH, MERGED_HEAVY_SMAX>(a, ch.x, a.p1, false, L.hi);
}

// One launch over processing positions [a.p0, a.p1) whose rows have at most `maxdeg` blocks. Lanes per row: 8 for
// scalar rows of up to 32 blocks (hex8: 27, in four sweeps; 32 rows per work-group, every lane busy in the gather of
// the 8-word scalar blocks: 0.46 ms vs 0.575 ms with 16 lanes on Poisson 128^3), 32 for vector rows of up to 32
// blocks, 64 for up to 64, and 128 for up to 128 (rows of 65 blocks: 128 lanes with one slot each measured 7.59 ms for
// the tet10 80^3 reduction against 7.95 ms with 64 lanes in two sweeps, whose second sweep is lane 0 alone; staging all
// 24 incidences in one pass with 8 loads in flight per lane then took it from 7.82 to 7.21 ms). The scatter kernel
// has since taken over the 3-field light rows and the tet10 vertex rows of up to 72 blocks; wider rows stay here.
template <int NF, int NNPE>
static int launch_lpr(const ReduceArgs& a, int maxdeg, hipStream_t s) {
  if (a.p1 <= a.p0) return 0;
  const int lpr = (NF == 1 && maxdeg <= 32) ? RED_NF1_LPR : (maxdeg <= 32 ? RED_NF3_LPR : (maxdeg <= 64 ? 64 : RED_HEAVY_LPR));
  const int sw = maxdeg <= lpr ? 1 : (maxdeg <= 2 * lpr ? 2 : 4);
  // more than one sweep on the narrow scalar rows, where wider rows move to the next lane count first, and on the
  // 128-lane rows (129..256 blocks: vertex rows of unstructured tet10 meshes); wider rows are refused
  if (maxdeg > sw * lpr || (sw == 4 && lpr != 8) || (sw > 1 && lpr > 16 && lpr != 128)) return -1;
  const int64_t npw = RED_THREADS / lpr;
  int64_t grid = (a.p1 - a.p0 + npw - 1) / npw;
  grid = (grid + 7) / 8 * 8;  // XCD remap needs a multiple of 8
  const unsigned g = (unsigned)grid;
  if constexpr (NF == 1) {
    if (lpr == 8) {
      if (sw == 1) reduce_rows_kernel<NF, NNPE, 8, 1><<<g, RED_THREADS, 0, s>>>(a);
      else if (sw == 2) reduce_rows_kernel<NF, NNPE, 8, 2><<<g, RED_THREADS, 0, s>>>(a);
      else reduce_rows_kernel<NF, NNPE, 8, 4><<<g, RED_THREADS, 0, s>>>(a);
      return 0;
    }
  }
  if (lpr == 16 && sw == 1) reduce_rows_kernel<NF, NNPE, 16, 1><<<g, RED_THREADS, 0, s>>>(a);
  else if (lpr == 16) reduce_rows_kernel<NF, NNPE, 16, 2><<<g, RED_THREADS, 0, s>>>(a);
  else if (lpr == 32) reduce_rows_kernel<NF, NNPE, 32, 1><<<g, RED_THREADS, 0, s>>>(a);
  else if (lpr == 64) reduce_rows_kernel<NF, NNPE, 64, 1><<<g, RED_THREADS, 0, s>>>(a);
  else if (sw == 1) reduce_rows_kernel<NF, NNPE, 128, 1, RED_HEAVY_CH><<<g, RED_THREADS, 0, s>>>(a);
  else reduce_rows_kernel<NF, NNPE, 128, 2, RED_HEAVY_CH><<<g, RED_THREADS, 0, s>>>(a);
  return 0;
}

// The processing order puts the light rows (at most CH incidences and 32 blocks) first and the heavy ones after
// (tet10: edge-midpoint rows, then vertex rows); each bucket gets the lane count that fits it, so light rows neither
// idle on a heavy row's lane count nor wait at a work-group barrier through its extra passes.
template <int NF, int NNPE>
static int launch_buckets(const ReduceArgs& a, hipStream_t s) {
  ReduceArgs lo = a, hi = a;
  lo.p1 = std::min(a.p1, std::max(a.p0, a.nlight));
  hi.p0 = std::max(a.p0, a.nlight);
  // Light rows take one pass, so their launch can skip the pass-count vote and its barriers: measured -2.6 % (Poisson
  // hex8 128^3) and -0.8 % (neo-Hookean hex8 96^3), but +3.4 % on the tet10 edge rows, which keep the vote.
  static_assert(CH >= 8, "light rows have at most 8 incidences and must take one pass");
  lo.onepass = NNPE == 8 ? 1 : 0;
  hi.onepass = 0;
  // direct slot addressing measured -1 % (scalar rows) and +0.8 % (vector rows) on the matrix reduction
  lo.padded = (NF == 1 && a.npad >= lo.p1) ? 1 : 0;
  hi.padded = 0;
  int rc = 0;
  if constexpr (NF == 1 && NNPE == 8) {
    if (lo.p1 > lo.p0) {
      if (a.maxdeg_light > 32) return -1;
      int64_t grid = (lo.p1 - lo.p0 + RED_THREADS / 8 - 1) / (RED_THREADS / 8);
      grid = (grid + 7) / 8 * 8;  // XCD remap needs a multiple of 8
      reduce_scalar_hex8_kernel<32><<<(unsigned)grid, RED_THREADS, 0, s>>>(lo);
    } else {
      rc = launch_lpr<NF, NNPE>(lo, a.maxdeg_light, s);
    }
  } else if constexpr (NF == 3) {
    if (NNPE == 10 && a.merged && a.nmerged > 0 && a.p0 == 0 && a.nlight < a.p1 && a.maxdeg_light <= 32 &&
        a.maxdeg <= MERGED_HEAVY_SMAX) {
      const int64_t grid = (a.nmerged + 7) / 8 * 8;  // XCD remap needs a multiple of 8
      reduce_merged_kernel<<<(unsigned)grid, RED_THREADS, 0, s>>>(a);
      return 0;
    }
    if (lo.p1 > lo.p0 && a.maxdeg_light <= 32) {
      int64_t grid = (lo.p1 - lo.p0 + RED_THREADS / 32 - 1) / (RED_THREADS / 32);
      grid = (grid + 7) / 8 * 8;  // XCD remap needs a multiple of 8
      reduce_scatter_kernel<NF, NNPE, 32, CH, 32><<<(unsigned)grid, RED_THREADS, 0, s>>>(lo);
    } else {
      rc = launch_lpr<NF, NNPE>(lo, a.maxdeg_light, s);
    }
  } else {
    rc = launch_lpr<NF, NNPE>(lo, a.maxdeg_light, s);
  }
  if (rc != 0 || hi.p1 <= hi.p0) return rc;
  if constexpr (NNPE == 10) {
    if (a.maxdeg <= HEAVY_SCATTER_SMAX) {
      int64_t grid = (hi.p1 - hi.p0 + RED_THREADS / 64 - 1) / (RED_THREADS / 64);
      grid = (grid + 7) / 8 * 8;  // XCD remap needs a multiple of 8
      reduce_scatter_kernel<NF, NNPE, 64, RED_HEAVY_CH, HEAVY_SCATTER_SMAX><<<(unsigned)grid, RED_THREADS, 0, s>>>(hi);
      return 0;
    }
  }
  return launch_lpr<NF, NNPE>(hi, a.maxdeg, s);
}

// Vector-only reduction (residual, actions, diagonals: no CSR values): one thread per row sums the row's incident
// element rows into R in ascending incidence order -- the order of the matrix kernel's residual sum, so the two give
// the same bits. Incidences go in batches of 8: indices first, then all their loads, then the FP64 adds.
template <int NF>
__global__ __launch_bounds__(256) void reduce_vector_kernel(ReduceArgs a) {
  const int64_t pos = a.p0 + xcd_chunk(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (pos >= a.p1) return;
  const int4 ri = a.rowinfo[pos];
  const int64_t n = ri.x;
  const int32_t c0 = ri.y, ninc = ri.w & 0xFFFF;
  double acc[NF];
#pragma unroll
  for (int i = 0; i < NF; ++i) acc[i] = 0.0;
  if (pos < a.npad) {
    // padded light row: its 8 slots sit at 8 pos, so their loads do not wait for the descriptor (slots past ninc hold a
    // valid incidence and are not summed)
    int32_t v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = a.vent[pos * 8 + t];
    float r[8][NF];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int i = 0; i < NF; ++i) r[t][i] = a.Rel[(int64_t)v[t] * NF + i];
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (t < ninc)
#pragma unroll
        for (int i = 0; i < NF; ++i) acc[i] += (double)r[t][i];
  } else {
    for (int t0 = 0; t0 < ninc; t0 += 8) {
      int32_t v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = t0 + t < ninc ? a.vent[c0 + t0 + t] : a.vent[c0];
      float r[8][NF];
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < NF; ++i) r[t][i] = a.Rel[(int64_t)v[t] * NF + i];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (t0 + t < ninc)
#pragma unroll
          for (int i = 0; i < NF; ++i) acc[i] += (double)r[t][i];
    }
  }
#pragma unroll
  for (int i = 0; i < NF; ++i) a.R[n * NF + i] = acc[i];
}

// Vector-only reduction with G lanes per row, for rows with many incidences (tet10 vertices: 24): lane l loads the
// incidences l, l + G, ... of each pass of CHV, all before any add, stages their element rows in LDS, and lane i < NF
// of the row then adds component i of incidences 0, 1, ... in ascending order: the order, and so the bits, of the
// thread-per-row kernel above and of the matrix kernel's residual sum.
template <int NF, int G, int CHV>
__global__ __launch_bounds__(256) void reduce_vector_group_kernel(ReduceArgs a) {
  constexpr int RPW = 256 / G, PER = CHV / G;
  static_assert(CHV % G == 0, "whole incidences per lane");
  __shared__ float s_r[RPW][CHV][NF];
  __shared__ int s_maxpass;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const int64_t pos = a.p0 + xcd_chunk(blockIdx.x, gridDim.x) * RPW + grp;
  const bool valid = pos < a.p1;
  const int4 ri = valid ? a.rowinfo[pos] : make_int4(0, 0, 0, 0);
  const int32_t c0 = ri.y, ninc = ri.w & 0xFFFF;
  if (threadIdx.x == 0) s_maxpass = 0;
  __syncthreads();
  if (lane == 0) atomicMax(&s_maxpass, (ninc + CHV - 1) / CHV);
  __syncthreads();
  const int maxpass = s_maxpass;
  double acc = 0.0;
  for (int pass = 0; pass < maxpass; ++pass) {
    const int base = pass * CHV;
    int32_t v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int t = base + lane + G * k;
      v[k] = t < ninc ? a.vent[c0 + t] : -1;
    }
    float r[PER][NF];
#pragma unroll
    for (int k = 0; k < PER; ++k)
#pragma unroll
      for (int i = 0; i < NF; ++i) r[k][i] = v[k] >= 0 ? a.Rel[(int64_t)v[k] * NF + i] : 0.0f;
#pragma unroll
    for (int k = 0; k < PER; ++k)
#pragma unroll
      for (int i = 0; i < NF; ++i) s_r[grp][lane + G * k][i] = r[k][i];
    __syncthreads();
    if (lane < NF) {
      const int cnt = min(CHV, ninc - base);
      for (int t = 0; t < cnt; ++t) acc += (double)s_r[grp][t][lane];
    }
    if (pass + 1 < maxpass) __syncthreads();
  }
  if (valid && lane < NF) a.R[(int64_t)ri.x * NF + lane] = acc;
}

int launch_reduce(int nf, const ReduceArgs& a, hipStream_t s) {
  if (a.p1 <= a.p0) return 0;
  if (!a.vals) {
    if (!a.R) return 0;
    if (a.nnpe == 10) {
      // light rows (edge midpoints, at most 8 incidences: one batch of loads) one thread each, then the vertex rows
      // (24 incidences) 16 lanes each, 32 incidences staged per pass (0.370 ms against 0.393 with 8 lanes and 24, and
      // 0.414 with 4, on tet10 80^3 x 6)
      ReduceArgs lo = a, hi = a;
      lo.p1 = std::min(a.p1, std::max(a.p0, a.nlight));
      hi.p0 = std::max(a.p0, a.nlight);
      if (nf != 1 && nf != 3) return -1;
      if (lo.p1 > lo.p0) {
        int64_t grid = (lo.p1 - lo.p0 + 255) / 256;
        grid = (grid + 7) / 8 * 8;  // XCD remap needs a multiple of 8
        if (nf == 1) reduce_vector_kernel<1><<<(unsigned)grid, 256, 0, s>>>(lo);
        else reduce_vector_kernel<3><<<(unsigned)grid, 256, 0, s>>>(lo);
      }
      if (hi.p1 > hi.p0) {
        constexpr int G = 16;
        int64_t grid = (hi.p1 - hi.p0 + 256 / G - 1) / (256 / G);
        grid = (grid + 7) / 8 * 8;
        if (nf == 1) reduce_vector_group_kernel<1, G, 32><<<(unsigned)grid, 256, 0, s>>>(hi);
        else reduce_vector_group_kernel<3, G, 32><<<(unsigned)grid, 256, 0, s>>>(hi);
      }
      return 0;
    }
    int64_t grid = (a.p1 - a.p0 + 255) / 256;
    grid = (grid + 7) / 8 * 8;  // XCD remap needs a multiple of 8
    if (nf == 1) reduce_vector_kernel<1><<<(unsigned)grid, 256, 0, s>>>(a);
    else if (nf == 2) reduce_vector_kernel<2><<<(unsigned)grid, 256, 0, s>>>(a);
    else if (nf == 3) reduce_vector_kernel<3><<<(unsigned)grid, 256, 0, s>>>(a);
    else return -1;
    return 0;
  }
  if (nf == 1) {
    if (a.nnpe == 4) return launch_buckets<1, 4>(a, s);
    if (a.nnpe == 8) return launch_buckets<1, 8>(a, s);
    if (a.nnpe == 10) return launch_buckets<1, 10>(a, s);
  } else if (nf == 2) {
    if (a.nnpe == 4) return launch_buckets<2, 4>(a, s);
  } else if (nf == 3) {
    if (a.nnpe == 8) return launch_buckets<3, 8>(a, s);
    if (a.nnpe == 10) return launch_buckets<3, 10>(a, s);
  }
  return -1;
}

// ------------------------------------------------------------------ DOF plumbing
__global__ void scatter_idx_kernel(int64_t n, const int64_t* idx, const double* src, double* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}
__global__ void gather_idx_kernel(int64_t n, const int64_t* idx, const double* src, double* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void copy_idx_kernel(int64_t n, const int64_t* idx, const double* src, double* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[idx[i]];
}
void launch_copy_indexed(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
  if (n > 0) copy_idx_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, idx, src, dst);
}

void launch_scatter_unknowns(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
  if (n > 0) scatter_idx_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, idx, src, dst);
}
void launch_gather_unknowns(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
  if (n > 0) gather_idx_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, idx, src, dst);
}

}  // namespace fecx
