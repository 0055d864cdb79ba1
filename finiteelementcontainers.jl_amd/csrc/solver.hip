// Linear algebra on the assembled CSR Jacobian: SpMV, diagonal, and Jacobi-preconditioned conjugate gradients in FP64.
//
// The reference solves each Newton step with Krylov.jl's CG on stiffness(asm) (src/Solvers.jl:128-153) on the CPU. Here
// the CSR values stay in HBM where the assembly wrote them. The SpMV reads the column indices from the block-CSR view
// (one column index per NF x NF block, NF consecutive columns per block), so the CSR colval array is never
// materialised on the device. All reductions are two-stage in a fixed order: deterministic and bit-reproducible. The
// CG scalars live in device memory; the host only checks convergence every few iterations.
#include "api_util.h"

namespace fecx {

constexpr int SV_THREADS = 256;
constexpr int SV_LANES = 16;  // lanes per CSR row in the SpMV

struct CsrView {
  int64_t nrows;            // unknowns = NF x free nodes
  const int64_t* bptr;      // free nodes + 1
  const int32_t* bcol;      // nnzb
  const double* vals;       // nnz
  const int64_t* rptr;      // dof-level CSR (a Dirichlet set that splits nodes): nrows + 1, or null
  const int32_t* rcol;      // its column indices
};

// work-group sum of one double per thread (fixed tree), written by thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = SV_THREADS / 2; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  return sh[0];
}

// y = A x; optionally partial[wg] = sum over the work-group's rows of w[row] * y[row]
template <int NF>
__global__ __launch_bounds__(SV_THREADS) void spmv_kernel(CsrView A, const double* x, double* y, const double* w,
                                                          double* partial, const int* done) {
  __shared__ double sh[SV_THREADS];
  if (done && *done) return;
  const int lane = threadIdx.x % SV_LANES;
  const int64_t row = (int64_t)blockIdx.x * (SV_THREADS / SV_LANES) + threadIdx.x / SV_LANES;
  double acc = 0.0;
  if (row < A.nrows && A.rptr) {
    for (int64_t j = A.rptr[row] + lane; j < A.rptr[row + 1]; j += SV_LANES) acc += A.vals[j] * x[A.rcol[j]];
  } else if (row < A.nrows) {
    const int64_t r = row / NF;
    const int i = (int)(row - r * NF);
    const int64_t rb = A.bptr[r];
    const int n = (int)(A.bptr[r + 1] - rb) * NF;
    const double* v = A.vals + NF * NF * rb + (int64_t)i * n;
    const int32_t* bc = A.bcol + rb;
    for (int j = lane; j < n; j += SV_LANES) {
      const int s = j / NF, k = j - s * NF;
      acc += v[j] * x[(int64_t)NF * bc[s] + k];
    }
  }
#pragma unroll
  for (int off = SV_LANES / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, SV_LANES);
  if (row < A.nrows && lane == 0) y[row] = acc;
  if (partial) {
    const double d = (row < A.nrows && lane == 0 && w) ? w[row] * acc : 0.0;
    const double s = block_sum(d, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// dinv[row] = 1 / A[row][row]; the diagonal block of free rank r is the slot s with bcol[rb + s] == r
template <int NF>
__global__ void diag_inv_kernel(CsrView A, double* dinv) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= A.nrows) return;
  if (A.rptr) {  // dof-level CSR: binary search for the diagonal column
    int64_t lo = A.rptr[row], hi = A.rptr[row + 1] - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (A.rcol[mid] < row) lo = mid + 1;
      else hi = mid;
    }
    const double d = (lo < A.rptr[row + 1] && A.rcol[lo] == row) ? A.vals[lo] : 0.0;
    dinv[row] = d != 0.0 ? 1.0 / d : 1.0;
    return;
  }
  const int64_t r = row / NF;
  const int i = (int)(row - r * NF);
  const int64_t rb = A.bptr[r];
  const int deg = (int)(A.bptr[r + 1] - rb);
  int lo = 0, hi = deg - 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (A.bcol[rb + mid] < r) lo = mid + 1;
    else hi = mid;
  }
  const double d = A.vals[NF * NF * rb + (int64_t)i * NF * deg + NF * lo + i];
  dinv[row] = d != 0.0 ? 1.0 / d : 1.0;
}

// ------------------------------------------------------------------ CG
// scalars: [0] rz, [1] pAp, [2] alpha, [3] beta, [4] rr, [5] tol^2, [6] iterations (as double)
struct CgState {
  int64_t n;
  double* x;
  double* r;
  double* z;
  double* p;
  double* Ap;
  const double* dinv;
  double* sc;
  int* done;
  double* part1;   // per work-group partials
  double* part2;
  double* part3;
  int64_t nparts;  // partial count of the vector kernels
};

// fixed-order sum of n partials by one work-group
__device__ __forceinline__ double sum_partials(const double* p, int64_t n, double* sh) {
  double a = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += SV_THREADS) a += p[k];
  return block_sum(a, sh);
}

// r = b - A x has been formed in r; z = dinv r; p = z; rz, rr, tol
__global__ __launch_bounds__(SV_THREADS) void cg_init_vec_kernel(CgState c) {
  __shared__ double sh[SV_THREADS];
  const int64_t i = (int64_t)blockIdx.x * SV_THREADS + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < c.n) {
    const double r = c.r[i], z = c.dinv ? c.dinv[i] * r : r;
    c.z[i] = z;
    c.p[i] = z;
    rz = r * z;
    rr = r * r;
  }
  const double a = block_sum(rz, sh);
  if (threadIdx.x == 0) c.part2[blockIdx.x] = a;
  __syncthreads();
  const double b = block_sum(rr, sh);
  if (threadIdx.x == 0) c.part3[blockIdx.x] = b;
}

__global__ __launch_bounds__(SV_THREADS) void cg_init_scalars_kernel(CgState c, double tol2) {
  __shared__ double sh[SV_THREADS];
  const double rz = sum_partials(c.part2, c.nparts, sh);
  __syncthreads();
  const double rr = sum_partials(c.part3, c.nparts, sh);
  if (threadIdx.x == 0) {
    c.sc[0] = rz;
    c.sc[4] = rr;
    c.sc[5] = tol2;
    c.sc[6] = 0.0;
    *c.done = rr <= tol2 ? 1 : 0;
  }
}

__global__ __launch_bounds__(SV_THREADS) void cg_alpha_kernel(CgState c, int64_t nspmv) {
  __shared__ double sh[SV_THREADS];
  if (*c.done) return;
  const double pap = sum_partials(c.part1, nspmv, sh);
  if (threadIdx.x == 0) {
    c.sc[1] = pap;
    c.sc[2] = pap != 0.0 ? c.sc[0] / pap : 0.0;
  }
}

// x += alpha p; r -= alpha Ap; z = dinv r; partials of r.z and r.r
__global__ __launch_bounds__(SV_THREADS) void cg_update_kernel(CgState c) {
  __shared__ double sh[SV_THREADS];
  if (*c.done) return;
  const double alpha = c.sc[2];
  const int64_t i = (int64_t)blockIdx.x * SV_THREADS + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < c.n) {
    c.x[i] += alpha * c.p[i];
    const double r = c.r[i] - alpha * c.Ap[i];
    c.r[i] = r;
    const double z = c.dinv ? c.dinv[i] * r : r;
    c.z[i] = z;
    rz = r * z;
    rr = r * r;
  }
  const double a = block_sum(rz, sh);
  if (threadIdx.x == 0) c.part2[blockIdx.x] = a;
  __syncthreads();
  const double b = block_sum(rr, sh);
  if (threadIdx.x == 0) c.part3[blockIdx.x] = b;
}

__global__ __launch_bounds__(SV_THREADS) void cg_beta_kernel(CgState c) {
  __shared__ double sh[SV_THREADS];
  if (*c.done) return;
  const double rz = sum_partials(c.part2, c.nparts, sh);
  __syncthreads();
  const double rr = sum_partials(c.part3, c.nparts, sh);
  if (threadIdx.x == 0) {
    c.sc[3] = c.sc[0] != 0.0 ? rz / c.sc[0] : 0.0;
    c.sc[0] = rz;
    c.sc[4] = rr;
    c.sc[6] += 1.0;
    if (rr <= c.sc[5]) *c.done = 1;
  }
}

__global__ __launch_bounds__(SV_THREADS) void cg_direction_kernel(CgState c) {
  if (*c.done) return;
  const double beta = c.sc[3];
  const int64_t i = (int64_t)blockIdx.x * SV_THREADS + threadIdx.x;
  if (i < c.n) c.p[i] = c.z[i] + beta * c.p[i];
}

// r = b - r (r holds A x on entry)
__global__ void residual_kernel(int64_t n, const double* b, double* r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

__global__ __launch_bounds__(SV_THREADS) void sumsq_kernel(int64_t n, const double* v, double* partial) {
  __shared__ double sh[SV_THREADS];
  const int64_t i = (int64_t)blockIdx.x * SV_THREADS + threadIdx.x;
  const double a = i < n ? v[i] * v[i] : 0.0;
  const double s = block_sum(a, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ __launch_bounds__(SV_THREADS) void sum_kernel(const double* partial, int64_t n, double* out) {
  __shared__ double sh[SV_THREADS];
  const double s = sum_partials(partial, n, sh);
  if (threadIdx.x == 0) *out = s;
}
__global__ void axpy_kernel(int64_t n, double a, const double* x, double* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

template <int NF>
static void spmv_launch(const CsrView& A, const double* x, double* y, const double* w, double* partial, const int* done,
                        hipStream_t s) {
  const unsigned g = (unsigned)((A.nrows + SV_THREADS / SV_LANES - 1) / (SV_THREADS / SV_LANES));
  spmv_kernel<NF><<<g, SV_THREADS, 0, s>>>(A, x, y, w, partial, done);
}

static void spmv(int nf, const CsrView& A, const double* x, double* y, const double* w, double* partial,
                 const int* done, hipStream_t s) {
  if (nf == 1) spmv_launch<1>(A, x, y, w, partial, done, s);
  else if (nf == 2) spmv_launch<2>(A, x, y, w, partial, done, s);
  else spmv_launch<3>(A, x, y, w, partial, done, s);
}

static int64_t spmv_groups(int64_t nrows) { return (nrows + SV_THREADS / SV_LANES - 1) / (SV_THREADS / SV_LANES); }

}  // namespace fecx

using namespace fecx_api;

namespace {

fecx::CsrView csr_view(fecx_ctx* ctx, const double* vals) {
  fecx::CsrView A{};
  A.nrows = (int64_t)ctx->hp.nfree * ctx->nf;
  A.bptr = get<int64_t>(ctx, "bptr");
  A.bcol = get<int32_t>(ctx, "bcol");
  A.vals = vals;
  if (ctx->partial) {
    A.nrows = (int64_t)ctx->unknown_dofs.size();
    A.rptr = get<int64_t>(ctx, "rptr");
    A.rcol = get<int32_t>(ctx, "rcol");
  }
  return A;
}

void check_solver_ready(fecx_ctx* ctx) {
  check_device_ctx(ctx);
  require(ctx->has_pattern, FECX_ERR_STATE, "fecx_build_pattern has not been called");
  require(ctx->nf >= 1 && ctx->nf <= 3, FECX_ERR_UNSUPPORTED, "nf must be 1, 2 or 3");
}

}  // namespace

extern "C" {

int fecx_csr_spmv(fecx_ctx* ctx, const double* vals, const double* x, double* y) {
  return guard(ctx, [&] {
    check_solver_ready(ctx);
    require(vals && x && y, FECX_ERR_ARG, "null array");
    const fecx::CsrView A = csr_view(ctx, vals);
    fecx::spmv(ctx->nf, A, x, y, nullptr, nullptr, nullptr, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_csr_diagonal_inverse(fecx_ctx* ctx, const double* vals, double* dinv) {
  return guard(ctx, [&] {
    check_solver_ready(ctx);
    require(vals && dinv, FECX_ERR_ARG, "null array");
    const fecx::CsrView A = csr_view(ctx, vals);
    if (A.nrows > 0) {
      const unsigned g = (unsigned)((A.nrows + 255) / 256);
      if (ctx->nf == 1) fecx::diag_inv_kernel<1><<<g, 256, 0, ctx->stream>>>(A, dinv);
      else if (ctx->nf == 2) fecx::diag_inv_kernel<2><<<g, 256, 0, ctx->stream>>>(A, dinv);
      else fecx::diag_inv_kernel<3><<<g, 256, 0, ctx->stream>>>(A, dinv);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_cg_solve(fecx_ctx* ctx, const double* vals, const double* b, double* x, double rtol, double atol, int maxit,
                  int precond, int* iters, double* resnorm) {
  return guard(ctx, [&] {
    check_solver_ready(ctx);
    require(vals && b && x && maxit >= 0 && rtol >= 0 && atol >= 0, FECX_ERR_ARG, "bad arguments");
    const fecx::CsrView A = csr_view(ctx, vals);
    const int64_t n = A.nrows;
    hipStream_t s = ctx->stream;
    fecx::CgState c{};
    c.n = n;
    c.x = x;
    c.r = (double*)dev_atleast(ctx, "cg_r", sizeof(double) * (n + 1));
    c.z = (double*)dev_atleast(ctx, "cg_z", sizeof(double) * (n + 1));
    c.p = (double*)dev_atleast(ctx, "cg_p", sizeof(double) * (n + 1));
    c.Ap = (double*)dev_atleast(ctx, "cg_ap", sizeof(double) * (n + 1));
    c.sc = (double*)dev_atleast(ctx, "cg_sc", sizeof(double) * 8);
    c.done = (int*)dev_atleast(ctx, "cg_done", sizeof(int));
    const int64_t ng = fecx::spmv_groups(n);
    c.nparts = (n + fecx::SV_THREADS - 1) / fecx::SV_THREADS;
    c.part1 = (double*)dev_atleast(ctx, "cg_part1", sizeof(double) * std::max<int64_t>(ng, 1));
    c.part2 = (double*)dev_atleast(ctx, "cg_part2", sizeof(double) * std::max<int64_t>(c.nparts, 1));
    c.part3 = (double*)dev_atleast(ctx, "cg_part3", sizeof(double) * std::max<int64_t>(c.nparts, 1));
    double* dinv = nullptr;
    if (precond) {
      dinv = (double*)dev_atleast(ctx, "cg_dinv", sizeof(double) * (n + 1));
      const unsigned g = (unsigned)((n + 255) / 256);
      if (n > 0) {
        if (ctx->nf == 1) fecx::diag_inv_kernel<1><<<g, 256, 0, s>>>(A, dinv);
        else if (ctx->nf == 2) fecx::diag_inv_kernel<2><<<g, 256, 0, s>>>(A, dinv);
        else fecx::diag_inv_kernel<3><<<g, 256, 0, s>>>(A, dinv);
      }
    }
    c.dinv = dinv;
    // ||b|| for the stopping rule ||r|| <= atol + rtol ||b|| (Krylov.jl's cg)
    const unsigned gv = (unsigned)std::max<int64_t>(c.nparts, 1);
    HIPCHK(hipMemcpyAsync(c.r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemsetAsync(c.done, 0, sizeof(int), s));
    fecx::CgState cb = c;
    cb.dinv = nullptr;  // b.b and r.r of r = b
    fecx::cg_init_vec_kernel<<<gv, fecx::SV_THREADS, 0, s>>>(cb);
    fecx::cg_init_scalars_kernel<<<1, fecx::SV_THREADS, 0, s>>>(cb, 0.0);
    double bb = 0.0;
    HIPCHK(hipMemcpyAsync(&bb, c.sc + 4, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double tol = atol + rtol * std::sqrt(bb);
    // r = b - A x
    fecx::spmv(ctx->nf, A, x, c.r, nullptr, nullptr, nullptr, s);
    fecx::residual_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, b, c.r);
    fecx::cg_init_vec_kernel<<<gv, fecx::SV_THREADS, 0, s>>>(c);
    fecx::cg_init_scalars_kernel<<<1, fecx::SV_THREADS, 0, s>>>(c, tol * tol);
    HIPCHK(hipGetLastError());
    int it = 0, hdone = 0;
    while (it < maxit) {
      const int batch = std::min(16, maxit - it);
      for (int k = 0; k < batch; ++k) {
        fecx::spmv(ctx->nf, A, c.p, c.Ap, c.p, c.part1, c.done, s);
        fecx::cg_alpha_kernel<<<1, fecx::SV_THREADS, 0, s>>>(c, ng);
        fecx::cg_update_kernel<<<gv, fecx::SV_THREADS, 0, s>>>(c);
        fecx::cg_beta_kernel<<<1, fecx::SV_THREADS, 0, s>>>(c);
        fecx::cg_direction_kernel<<<gv, fecx::SV_THREADS, 0, s>>>(c);
      }
      HIPCHK(hipGetLastError());
      it += batch;
      HIPCHK(hipMemcpyAsync(&hdone, c.done, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (hdone) break;
    }
    double sc[8];
    HIPCHK(hipMemcpy(sc, c.sc, sizeof(sc), hipMemcpyDeviceToHost));
    if (iters) *iters = (int)sc[6];
    if (resnorm) *resnorm = std::sqrt(sc[4]);
  });
}

int fecx_vector_norm(fecx_ctx* ctx, const double* v, int64_t n, double* out) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(n >= 0 && (n == 0 || v) && out, FECX_ERR_ARG, "bad arguments");
    const int64_t np = std::max<int64_t>((n + fecx::SV_THREADS - 1) / fecx::SV_THREADS, 1);
    double* part = (double*)dev_atleast(ctx, "vn_part", sizeof(double) * np);
    double* res = (double*)dev_atleast(ctx, "vn_res", sizeof(double));
    fecx::sumsq_kernel<<<(unsigned)np, fecx::SV_THREADS, 0, ctx->stream>>>(n, v, part);
    fecx::sum_kernel<<<1, fecx::SV_THREADS, 0, ctx->stream>>>(part, np, res);
    HIPCHK(hipGetLastError());
    double h = 0.0;
    HIPCHK(hipMemcpyAsync(&h, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = std::sqrt(h);
  });
}

int fecx_vector_axpy(fecx_ctx* ctx, int64_t n, double a, const double* x, double* y) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(n >= 0 && (n == 0 || (x && y)), FECX_ERR_ARG, "bad arguments");
    if (n > 0) fecx::axpy_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(n, a, x, y);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
