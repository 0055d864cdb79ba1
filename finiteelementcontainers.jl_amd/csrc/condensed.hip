// Condensed mode: the constraint adjustments of an assembler built with use_condensed = true.
//
// Reference: src/assemblers/Utils.jl:53-167. A condensed assembler keeps every DOF in its sparsity pattern (the
// pattern is never reduced, src/assemblers/SparseMatrixAssembler.jl:246-251) and its accessors adjust the assembled
// results with constraint_storage c (1 at Dirichlet dofs, 0 elsewhere, SparseMatrixAssembler.jl:93-94):
//   stiffness / mass   compressed line k (a CSR row; a CSC column): nz = (1 - c[k]) nz, and its diagonal entry
//                      + penalty c[k], penalty = penalty_scale tr(A) / n, the trace taken before the adjustment
//   residual           b = (1 - c) b
//   hvp(asm, v)        Av = (1 - c) Av + c v
// Only the lines of constrained dofs change (a factor 1 - c = 1 leaves an entry as it is), so the kernels visit the
// Dirichlet list alone. The trace is a fixed-order two-stage FP64 sum, so the penalty is the same from run to run.
// For the symmetric tangents of every registered physics the CSR values are symmetric bit for bit, so the CSC
// form (colptr, rowval) is the CSR (rowptr, colval) and zeroing CSC column k touches the same values as zeroing CSR
// row k: one kernel serves both.
#include <cmath>

#include "api_util.h"

namespace fecx {

constexpr int CD_THREADS = 256;

__device__ __forceinline__ double cd_block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = CD_THREADS / 2; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  return sh[0];
}

// offset of row (free rank r, component i)'s diagonal entry within its line of NF deg values: the diagonal block of
// free rank r is the slot s with bcol[rb + s] == r
template <int NF>
__device__ __forceinline__ int diag_in_line(const int64_t* bptr, const int32_t* bcol, int64_t r, int i) {
  const int64_t rb = bptr[r];
  const int deg = (int)(bptr[r + 1] - rb);
  int lo = 0, hi = deg - 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (bcol[rb + mid] < r) lo = mid + 1;
    else hi = mid;
  }
  return NF * lo + i;
}
template <int NF>
__device__ __forceinline__ int64_t diag_pos(const int64_t* bptr, const int32_t* bcol, int64_t r, int i) {
  const int64_t rb = bptr[r];
  return NF * NF * rb + (int64_t)i * NF * (bptr[r + 1] - rb) + diag_in_line<NF>(bptr, bcol, r, i);
}

// stage 1: partial[wg] = sum of the diagonal entries of the work-group's rows (in row order within a fixed tree)
template <int NF>
__global__ __launch_bounds__(CD_THREADS) void trace_partial_kernel(int64_t nrows, const int64_t* bptr,
                                                                     const int32_t* bcol, const double* vals,
                                                                     double* partial) {
  __shared__ double sh[CD_THREADS];
  const int64_t row = (int64_t)blockIdx.x * CD_THREADS + threadIdx.x;
  double d = 0.0;
  if (row < nrows) d = vals[diag_pos<NF>(bptr, bcol, row / NF, (int)(row % NF))];
  const double s = cd_block_sum(d, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// stage 2: penalty = scale * (sum of the partials) / n
__global__ __launch_bounds__(CD_THREADS) void penalty_kernel(const double* partial, int64_t np, double scale,
                                                             int64_t n, double* penalty) {
  __shared__ double sh[CD_THREADS];
  double a = 0.0;
  for (int64_t k = threadIdx.x; k < np; k += CD_THREADS) a += partial[k];
  const double tr = cd_block_sum(a, sh);
  if (threadIdx.x == 0) *penalty = scale * tr / (double)n;
}

// one 64-lane group per constrained dof d: its line nz = (1 - 1) nz, diagonal + penalty * 1
template <int NF>
__global__ __launch_bounds__(CD_THREADS) void constrain_lines_kernel(int64_t ncon, const int64_t* con,
                                                                     const int64_t* bptr, const int32_t* bcol,
                                                                     double* vals, const double* penalty) {
  const int64_t k = (int64_t)blockIdx.x * (CD_THREADS / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (k >= ncon) return;
  const int64_t d = con[k];  // 0-based dof: free rank d / NF (every node is free in condensed mode), component d % NF
  const int64_t r = d / NF;
  const int i = (int)(d % NF);
  const int64_t rb = bptr[r];
  const int n = (int)(bptr[r + 1] - rb) * NF;
  double* line = vals + NF * NF * rb + (int64_t)i * n;
  const int jd = diag_in_line<NF>(bptr, bcol, r, i);
  const double c = 1.0, pen = *penalty;
  for (int j = lane; j < n; j += 64) line[j] = j == jd ? (1.0 - c) * line[j] + pen * c : (1.0 - c) * line[j];
}

__global__ void constrain_vector_kernel(int64_t ncon, const int64_t* con, double* b, const double* v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncon) return;
  const int64_t d = con[k];
  const double c = 1.0;
  b[d] = v ? (1.0 - c) * b[d] + c * v[d] : (1.0 - c) * b[d];
}

}  // namespace fecx

using namespace fecx_api;

namespace {

void check_condensed(fecx_ctx* ctx) {
  check_device_ctx(ctx);
  require(ctx->has_pattern, FECX_ERR_STATE, "fecx_build_pattern has not been called");
  require(ctx->condensed, FECX_ERR_STATE, "the context is not in condensed mode (fecx_set_condensed)");
}

}  // namespace

extern "C" {

int fecx_adjust_matrix_for_constraints(fecx_ctx* ctx, double* vals, double penalty_scale) {
  return guard(ctx, [&] {
    check_condensed(ctx);
    require(vals != nullptr || ctx->nnz == 0, FECX_ERR_ARG, "vals is NULL");
    require(std::isfinite(penalty_scale), FECX_ERR_ARG, "penalty_scale must be finite");
    const int nf = ctx->nf;
    const int64_t nrows = ctx->hp.nfree * nf;
    const int64_t ncon = (int64_t)ctx->dirichlet.size();
    if (ncon == 0 || nrows == 0) return;
    const int64_t* bptr = get<int64_t>(ctx, "bptr");
    const int32_t* bcol = get<int32_t>(ctx, "bcol");
    const int64_t* con = get<int64_t>(ctx, "dirichlet_idx");
    const int64_t np = (nrows + fecx::CD_THREADS - 1) / fecx::CD_THREADS;
    double* part = (double*)dev_atleast(ctx, "cd_part", sizeof(double) * np);
    double* pen = (double*)dev_atleast(ctx, "cd_penalty", sizeof(double));
    const unsigned gl = (unsigned)((ncon + fecx::CD_THREADS / 64 - 1) / (fecx::CD_THREADS / 64));
    hipStream_t s = ctx->stream;
    if (nf == 1) {
      fecx::trace_partial_kernel<1><<<(unsigned)np, fecx::CD_THREADS, 0, s>>>(nrows, bptr, bcol, vals, part);
      fecx::penalty_kernel<<<1, fecx::CD_THREADS, 0, s>>>(part, np, penalty_scale, nrows, pen);
      fecx::constrain_lines_kernel<1><<<gl, fecx::CD_THREADS, 0, s>>>(ncon, con, bptr, bcol, vals, pen);
    } else if (nf == 2) {
      fecx::trace_partial_kernel<2><<<(unsigned)np, fecx::CD_THREADS, 0, s>>>(nrows, bptr, bcol, vals, part);
      fecx::penalty_kernel<<<1, fecx::CD_THREADS, 0, s>>>(part, np, penalty_scale, nrows, pen);
      fecx::constrain_lines_kernel<2><<<gl, fecx::CD_THREADS, 0, s>>>(ncon, con, bptr, bcol, vals, pen);
    } else {
      fecx::trace_partial_kernel<3><<<(unsigned)np, fecx::CD_THREADS, 0, s>>>(nrows, bptr, bcol, vals, part);
      fecx::penalty_kernel<<<1, fecx::CD_THREADS, 0, s>>>(part, np, penalty_scale, nrows, pen);
      fecx::constrain_lines_kernel<3><<<gl, fecx::CD_THREADS, 0, s>>>(ncon, con, bptr, bcol, vals, pen);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
  });
}

int fecx_adjust_vector_for_constraints(fecx_ctx* ctx, double* b) {
  return guard(ctx, [&] {
    check_condensed(ctx);
    require(b != nullptr, FECX_ERR_ARG, "b is NULL");
    const int64_t ncon = (int64_t)ctx->dirichlet.size();
    if (ncon > 0)
      fecx::constrain_vector_kernel<<<(unsigned)((ncon + 255) / 256), 256, 0, ctx->stream>>>(
          ncon, get<int64_t>(ctx, "dirichlet_idx"), b, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_adjust_action_for_constraints(fecx_ctx* ctx, double* Av, const double* v) {
  return guard(ctx, [&] {
    check_condensed(ctx);
    require(Av != nullptr && v != nullptr, FECX_ERR_ARG, "Av or v is NULL");
    const int64_t ncon = (int64_t)ctx->dirichlet.size();
    if (ncon > 0)
      fecx::constrain_vector_kernel<<<(unsigned)((ncon + 255) / 256), 256, 0, ctx->stream>>>(
          ncon, get<int64_t>(ctx, "dirichlet_idx"), Av, v);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
