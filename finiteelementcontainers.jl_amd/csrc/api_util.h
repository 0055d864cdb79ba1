// Helpers shared by the C-ABI translation units: error handling, device buffers owned by the context.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "fecx_internal.h"

namespace fecx_api {

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                      \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess) throw fecx_api::ApiError(FECX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class F>
inline int guard(fecx_ctx* ctx, F f) {
  if (!ctx) return FECX_ERR_ARG;
  ctx->err.clear();
  try {
    if (ctx->device >= 0) HIPCHK(hipSetDevice(ctx->device));
    f();
    return FECX_OK;
  } catch (const ApiError& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return FECX_ERR_ARG;
  }
}

inline void require(bool c, int code, const std::string& msg) {
  if (!c) throw ApiError(code, msg);
}

inline void check_device_ctx(fecx_ctx* ctx) {
  require(ctx->device >= 0, FECX_ERR_STATE, "host-only context (device -1) cannot run kernels");
}

// device buffers; a host-only context (device -1) allocates nothing
inline void* dev(fecx_ctx* ctx, const std::string& name, size_t bytes) {
  if (ctx->device < 0) return nullptr;
  DevBuf& b = ctx->buf[name];
  if (b.bytes != bytes) {
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    if (bytes) HIPCHK(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
  }
  return b.p;
}

// a buffer of at least `bytes`, grown (contents dropped) when too small
inline void* dev_atleast(fecx_ctx* ctx, const std::string& name, size_t bytes) {
  if (ctx->device < 0) return nullptr;
  DevBuf& b = ctx->buf[name];
  if (b.bytes < bytes) return dev(ctx, name, bytes);
  return b.p;
}

template <class T>
T* upload(fecx_ctx* ctx, const std::string& name, const std::vector<T>& v) {
  T* p = (T*)dev(ctx, name, v.size() * sizeof(T));
  if (p && !v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

inline size_t get_bytes(fecx_ctx* ctx, const std::string& name) {
  auto it = ctx->buf.find(name);
  return it == ctx->buf.end() ? 0 : it->second.bytes;
}

template <class T>
T* get(fecx_ctx* ctx, const std::string& name) {
  auto it = ctx->buf.find(name);
  return it == ctx->buf.end() ? nullptr : (T*)it->second.p;
}

}  // namespace fecx_api
