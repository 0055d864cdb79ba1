// HBM bandwidth ceilings of the device, measured with hand-written streaming kernels: the attainable figures next to
// the 8 TB/s spec that the assembly kernels' rooflines are priced against. No counterpart in the reference.
//   read   sum of every 16-byte word of a buffer (one FP64 add chain per lane, written once per work-group)
//   write  fill of a buffer with 16-byte stores
//   copy   16-byte loads and stores from one half of a buffer to the other
// Grid-stride loops over 16-byte words, 256-thread work-groups, 8 per CU; each lane keeps 4 loads in flight.
#include <algorithm>

#include "api_util.h"

namespace fecx {

typedef double f64x2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void bw_read_kernel(const f64x2* __restrict__ src, int64_t n, double* out) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  double s = 0.0;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const f64x2 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    s += a.x + a.y + b.x + b.y + c.x + c.y + d.x + d.y;
  }
  for (; i < n; i += stride) s += src[i].x + src[i].y;
  if (s == 12345.678) out[blockIdx.x] = s;  // keeps the loads; practically never taken
}

__global__ __launch_bounds__(256) void bw_write_kernel(f64x2* __restrict__ dst, int64_t n, double v) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  const f64x2 w = {v, v};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = w;
}

__global__ __launch_bounds__(256) void bw_copy_kernel(const f64x2* __restrict__ src, f64x2* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const f64x2 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a; dst[i + stride] = b; dst[i + 2 * stride] = c; dst[i + 3 * stride] = d;
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace fecx

using namespace fecx_api;

extern "C" {

int fecx_measure_bandwidth(fecx_ctx* ctx, int op, void* buf, int64_t bytes, int reps, double* ms) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(op >= 0 && op <= 2 && buf && ms && reps > 0 && bytes >= 4096 && bytes % 32 == 0, FECX_ERR_ARG,
            "op 0..2, a device buffer of a multiple of 32 bytes (>= 4096) and reps > 0");
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned grid = (unsigned)std::max(1, cus * 8);
    hipStream_t s = ctx->stream;
    fecx::f64x2* b = (fecx::f64x2*)buf;
    const int64_t n = bytes / 16, half = n / 2;
    double* sink = (double*)dev_atleast(ctx, "bw_sink", sizeof(double) * grid);
    auto launch = [&] {
      if (op == 0) fecx::bw_read_kernel<<<grid, 256, 0, s>>>(b, n, sink);
      else if (op == 1) fecx::bw_write_kernel<<<grid, 256, 0, s>>>(b, n, 1.0);
      else fecx::bw_copy_kernel<<<grid, 256, 0, s>>>(b, b + half, half);
    };
    launch();  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    *ms = t / reps;
  });
}

}  // extern "C"
