// Device-resident scratch for the allocation-free assembly entry point (fecx_assemble_vector_matrix_noalloc).
//
// The two-phase assembly keeps its element-private FP32 outputs (element matrices and rows) in context-owned device
// buffers. Callers that cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc -- get the same kernels with those outputs in this file-scope array instead. Each GPU the
// library runs on holds its own copy of the array (NOALLOC_SCRATCH_BYTES) until the process ends. In host code the
// array's name is only a stand-in for its device copy: the device address of each copy is looked up once per device
// with hipGetSymbolAddress (at fecx_create, with that device current) and kept per device; only that address is
// passed to kernels.
#include <mutex>

#include "api_util.h"

namespace fecx {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return FECX_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return FECX_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return FECX_ERR_HIP;  // the caller makes `device` current
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return FECX_ERR_HIP;
  g_scratch_ptr[device] = p;
  return FECX_OK;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace fecx
