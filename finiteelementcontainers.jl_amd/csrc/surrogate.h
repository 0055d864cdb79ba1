// Kernel arguments and launchers of the learned-surrogate kernels (mlp.hip), shared with the host side (surrogate.cpp).
#pragma once

#include <type_traits>

#include "fecx_internal.h"
#include "mlp_device.h"

namespace fecx {

struct MlpPack {  // writable views of the packed weights
  bf16x8* wf;
  bf16x8* wb;
  float* bias;
};

struct TrainArgs {
  MlpDev m;
  int64_t n, npad;       // samples in the batch, padded count (multiple of 256)
  const float* X;        // raw inputs, rows idx[s] (or s)
  const float* T;        // raw targets
  const int32_t* idx;    // optional row map of the batch
  __bf16* hbuf;          // h_l, l < NL: hrows(l) x npad (feature-major)
  __bf16* dzbuf;         // dz_l, 1 <= l <= NL: P[l] x npad
  __bf16* gpbuf;         // GELU'(z_l), 1 <= l < NL: [tile][register][lane]
  double* loss_part;     // npad / 32
};

struct WGradArgs {
  const __bf16* hbuf;
  const __bf16* dzbuf;
  int64_t npad;
  int nchunk, chunk_len;  // nchunk * chunk_len == npad
  float* part;            // nchunk x nparams
};

struct UpdateArgs {
  int64_t np;
  float* w;               // canonical FP32 parameters
  const float* part;      // chunk partial gradients, or nullptr (repack only)
  int nchunk;
  float* grad;            // optional: summed gradient out
  int adam, repack;
  float* m1;
  float* m2;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
  MlpPack pk;
};

struct MatBatchArgs {
  int64_t n;
  const float* in;        // batch mode: n x IN_W raw inputs
  float* out;             // batch mode: n x OUT_W raw outputs
  const float* hq;        // assembly mode: n x 9 quadrature-point gradients (element storage order)
  const int32_t* eperm;   // storage position -> element id (state index)
  int nq;
  const float* st_old;
  float* st_new;
  float* qpmat;           // n x (9 + NC)
  MlpDev m;
  MatParams mp;
  float tol;
  unsigned long long* nfallback;
  int* flag;
  uint8_t* accept;        // per point: 1 = network output used, 0 = plain model (assembly mode)
  int freeze;             // 1: reuse `accept` from the previous pass instead of checking again
};

void mlp_launch_train(int arch, const TrainArgs& a, hipStream_t s);
void mlp_launch_wgrad(int arch, const WGradArgs& a, hipStream_t s);
void mlp_launch_update(int arch, const UpdateArgs& a, hipStream_t s);
void mlp_launch_loss_sum(const double* part, int64_t n, double* out, hipStream_t s);
void mlp_launch_gather(int64_t n, int w, const int32_t* idx, const float* src, float* dst, hipStream_t s);
void mlp_launch_forward(int arch, const MatBatchArgs& a, hipStream_t s);
void mlp_launch_material(int arch, bool surrogate, const MatBatchArgs& a, hipStream_t s);
void mlp_launch_j2_paths(int64_t npaths, int steps, float max_norm, uint64_t seed, const MatParams& mp, float* X,
                         float* T, hipStream_t s);
void mlp_launch_column_stats(const float* X, int64_t n, int w, int ncol, float* shift, float* scale, hipStream_t s);
void mlp_launch_j2_features(const float* in, int64_t n, float* x, hipStream_t s);

// assembly hook (surrogate materials): material pass over all quadrature points, then the fallback count
void surrogate_qp_pass(fecx_ctx* ctx, const float* hq, float* qpmat, hipStream_t s);
void surrogate_qp_finish(fecx_ctx* ctx);

// shapes on the host side
struct ArchInfo {
  int nl;
  int d[5], p[5];
  int in_w, nfeat, out_w;
  int64_t nparams, wf_vecs, wb_vecs;
  int nbias;
  int wgrad_tiles;
};
ArchInfo arch_info(int arch);

}  // namespace fecx
