// Reduction kernels (phase 2 of assembly): FP64 sums of the FP32 element-private outputs.
//
// Matrix: the reference reduces its element-private COO storage with SparseArrays.sparse! over the sorted
// permutation on every stiffness() call (src/assemblers/SparsityPatterns.jl:301-329). Here each thread owns one
// NF x NF block of the CSR pattern and sums, in FP64 and in ascending element order (the order sparse! uses),
// the element blocks listed for it by the inverse COO->CSR map, then writes the block's NF rows x NF columns
// with plain coalesced stores: no atomics, no zero fill, every CSR value written exactly once.
// Vector: each thread owns one node and sums its incident element rows (the atomic scatter of
// src/assemblers/Assemblers.jl:72-87 restated as an owner-computes gather).
#include "fecx_internal.h"

namespace fecx {

constexpr int RED_THREADS = 256;
constexpr int ROWS_PER_WG = 16;

// Work-groups are dealt to the 8 XCDs round-robin; remap so that each XCD walks a contiguous range of
// block rows, which keeps the element blocks shared by neighbouring rows in that XCD's L2.
__device__ __forceinline__ int64_t xcd_chunk(int64_t wg, int64_t nwg) {
  const int64_t per = (nwg + 7) / 8;
  const int64_t x = wg % 8, i = wg / 8;
  return x * per + i;
}

template <int NF>
__global__ __launch_bounds__(RED_THREADS) void reduce_matrix_kernel(ReduceArgs a) {
  __shared__ int64_t sptr[ROWS_PER_WG + 1];
  const int64_t nwg = gridDim.x;
  const int64_t chunk = xcd_chunk(blockIdx.x, nwg);
  const int64_t r0 = chunk * ROWS_PER_WG;
  if (r0 >= a.nfree) return;  // whole work-group exits together
  const int64_t r1 = min(r0 + ROWS_PER_WG, a.nfree);
  const int nr = (int)(r1 - r0);
  if (threadIdx.x <= nr) sptr[threadIdx.x] = a.bptr[r0 + threadIdx.x];
  __syncthreads();
  const int64_t B0 = sptr[0], B1 = sptr[nr];
  for (int64_t b = B0 + threadIdx.x; b < B1; b += RED_THREADS) {
    int lo = 0, hi = nr - 1;  // row r with sptr[r] <= b < sptr[r+1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sptr[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int64_t rb = sptr[lo], deg = sptr[lo + 1] - rb, s = b - rb;
    double acc[NF * NF];
#pragma unroll
    for (int c = 0; c < NF * NF; ++c) acc[c] = 0.0;
    const int64_t c0 = a.boff[b], c1 = a.boff[b + 1];
    for (int64_t c = c0; c < c1; ++c) {
      const uint32_t ent = a.bent[c];
      const float* blk = a.Kel + (int64_t)(ent & 0x7fffffffu) * (NF * NF);
      if (ent >> 31) {
#pragma unroll
        for (int i = 0; i < NF; ++i)
#pragma unroll
          for (int k = 0; k < NF; ++k) acc[i * NF + k] += (double)blk[k * NF + i];
      } else {
#pragma unroll
        for (int c2 = 0; c2 < NF * NF; ++c2) acc[c2] += (double)blk[c2];
      }
    }
    const int64_t base = (int64_t)NF * NF * rb + (int64_t)NF * s;
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
      for (int k = 0; k < NF; ++k) a.vals[base + (int64_t)i * NF * deg + k] = acc[i * NF + k];
  }
}

template <int NF>
__global__ __launch_bounds__(RED_THREADS) void reduce_vector_kernel(ReduceArgs a) {
  const int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x;
  if (n >= a.nn) return;
  double acc[NF];
#pragma unroll
  for (int i = 0; i < NF; ++i) acc[i] = 0.0;
  const int32_t c0 = a.voff[n], c1 = a.voff[n + 1];
  for (int32_t c = c0; c < c1; ++c) {
    const int64_t slot = (int64_t)a.vent[c] * NF;
#pragma unroll
    for (int i = 0; i < NF; ++i) acc[i] += (double)a.Rel[slot + i];
  }
#pragma unroll
  for (int i = 0; i < NF; ++i) a.R[n * NF + i] = acc[i];
}

int launch_reduce_matrix(int nf, const ReduceArgs& a, hipStream_t s) {
  const unsigned grid = (unsigned)((a.nfree + ROWS_PER_WG - 1) / ROWS_PER_WG);
  if (grid == 0) return 0;
  // the XCD remap covers chunks 0 .. 8*ceil(grid/8)-1; pad the grid so every chunk is visited
  const unsigned padded = (unsigned)(((grid + 7) / 8) * 8);
  if (nf == 1) reduce_matrix_kernel<1><<<padded, RED_THREADS, 0, s>>>(a);
  else if (nf == 2) reduce_matrix_kernel<2><<<padded, RED_THREADS, 0, s>>>(a);
  else if (nf == 3) reduce_matrix_kernel<3><<<padded, RED_THREADS, 0, s>>>(a);
  else return -1;
  return 0;
}

int launch_reduce_vector(int nf, const ReduceArgs& a, hipStream_t s) {
  const unsigned grid = (unsigned)((a.nn + RED_THREADS - 1) / RED_THREADS);
  if (grid == 0) return 0;
  if (nf == 1) reduce_vector_kernel<1><<<grid, RED_THREADS, 0, s>>>(a);
  else if (nf == 2) reduce_vector_kernel<2><<<grid, RED_THREADS, 0, s>>>(a);
  else if (nf == 3) reduce_vector_kernel<3><<<grid, RED_THREADS, 0, s>>>(a);
  else return -1;
  return 0;
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

void launch_scatter_unknowns(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
  if (n > 0) scatter_idx_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, idx, src, dst);
}
void launch_gather_unknowns(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
  if (n > 0) gather_idx_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, idx, src, dst);
}

}  // namespace fecx
