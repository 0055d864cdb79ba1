// C-ABI entry points of the learned constitutive surrogate (declared in include/fecx.h). No counterpart in the
// reference: its materials are AD-differentiated energies evaluated inside the element loop
// (test/mechanics/TestMechanicsCommon.jl:140-222).
#include <cmath>
#include <cstring>
#include <numeric>

#include "api_util.h"
#include "surrogate.h"

using namespace fecx_api;

namespace {

// SplitMix64 stream, the same generator the Python side uses (finiteelementcontainers.jl_amd/rng.py)
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

void require_net(fecx_ctx* ctx) {
  check_device_ctx(ctx);
  require(ctx->sg_arch >= 0, FECX_ERR_STATE, "fecx_surrogate_init has not been called");
}

fecx::MlpDev mlp_dev(fecx_ctx* ctx) {
  fecx::MlpDev m{};
  m.wf = get<fecx::bf16x8>(ctx, "sg_wf");
  m.wb = get<fecx::bf16x8>(ctx, "sg_wb");
  m.bias = get<float>(ctx, "sg_bias");
  m.in_shift = get<float>(ctx, "sg_in_shift");
  m.in_scale = get<float>(ctx, "sg_in_scale");
  m.out_shift = get<float>(ctx, "sg_out_shift");
  m.out_scale = get<float>(ctx, "sg_out_scale");
  return m;
}

fecx::UpdateArgs update_args(fecx_ctx* ctx) {
  const fecx::ArchInfo ai = fecx::arch_info(ctx->sg_arch);
  fecx::UpdateArgs u{};
  u.np = ai.nparams;
  u.w = get<float>(ctx, "sg_w");
  u.m1 = get<float>(ctx, "sg_m1");
  u.m2 = get<float>(ctx, "sg_m2");
  u.pk.wf = get<fecx::bf16x8>(ctx, "sg_wf");
  u.pk.wb = get<fecx::bf16x8>(ctx, "sg_wb");
  u.pk.bias = get<float>(ctx, "sg_bias");
  u.b1 = 0.9f;
  u.b2 = 0.999f;
  u.eps = 1e-8f;
  return u;
}

void repack(fecx_ctx* ctx) {
  fecx::UpdateArgs u = update_args(ctx);
  u.repack = 1;
  fecx::mlp_launch_update(ctx->sg_arch, u, ctx->stream);
  HIPCHK(hipGetLastError());
}

int64_t npad_of(int64_t n) { return (n + 255) / 256 * 256; }

// One forward + backward pass over n samples (rows idx[0..n) of X/T, or 0..n); the summed gradient goes to `grad`
// (optional) and, with adam, into an Adam step. The mean loss is written to the device double *loss.
void grad_pass(fecx_ctx* ctx, const float* X, const float* T, const int32_t* idx, int64_t n, float* grad, bool adam,
               float lr, double* loss) {
  const int arch = ctx->sg_arch;
  const fecx::ArchInfo ai = fecx::arch_info(arch);
  const int64_t npad = npad_of(n);
  int64_t hrows = 32, dzrows = 0, gprows = 0;
  for (int l = 1; l < ai.nl; ++l) { hrows += ai.p[l]; gprows += ai.p[l]; }
  for (int l = 1; l <= ai.nl; ++l) dzrows += ai.p[l];
  const int chunk_len = 256;
  const int nchunk = (int)(npad / chunk_len);
  fecx::TrainArgs ta{};
  ta.m = mlp_dev(ctx);
  ta.n = n;
  ta.npad = npad;
  ta.X = X;
  ta.T = T;
  ta.idx = idx;
  ta.hbuf = (__bf16*)dev_atleast(ctx, "sg_h", sizeof(__bf16) * hrows * npad);
  ta.dzbuf = (__bf16*)dev_atleast(ctx, "sg_dz", sizeof(__bf16) * dzrows * npad);
  ta.gpbuf = (__bf16*)dev_atleast(ctx, "sg_gp", sizeof(__bf16) * gprows * npad);
  ta.loss_part = (double*)dev_atleast(ctx, "sg_lossp", sizeof(double) * (npad / 32));
  fecx::mlp_launch_train(arch, ta, ctx->stream);
  HIPCHK(hipGetLastError());
  fecx::WGradArgs wa{};
  wa.hbuf = ta.hbuf;
  wa.dzbuf = ta.dzbuf;
  wa.npad = npad;
  wa.nchunk = nchunk;
  wa.chunk_len = chunk_len;
  wa.part = (float*)dev_atleast(ctx, "sg_part", sizeof(float) * nchunk * ai.nparams);
  fecx::mlp_launch_wgrad(arch, wa, ctx->stream);
  HIPCHK(hipGetLastError());
  fecx::UpdateArgs u = update_args(ctx);
  u.part = wa.part;
  u.nchunk = nchunk;
  u.grad = grad;
  if (adam) {
    ctx->sg_step += 1;
    u.adam = 1;
    u.repack = 1;
    u.lr = lr;
    u.c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)ctx->sg_step)));
    u.c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)ctx->sg_step)));
  }
  fecx::mlp_launch_update(arch, u, ctx->stream);
  HIPCHK(hipGetLastError());
  fecx::mlp_launch_loss_sum(ta.loss_part, npad / 32, loss, ctx->stream);
  HIPCHK(hipGetLastError());
}

fecx::MatBatchArgs batch_args(fecx_ctx* ctx) {
  fecx::MatBatchArgs a{};
  a.m = mlp_dev(ctx);
  a.mp = ctx->mp;
  a.tol = ctx->sg_tol;
  a.nq = ctx->nq;
  a.nfallback = (unsigned long long*)dev_atleast(ctx, "sg_nfb", sizeof(unsigned long long));
  return a;
}

}  // namespace

namespace fecx {

void surrogate_qp_pass(fecx_ctx* ctx, const float* hq, float* qpmat, hipStream_t s) {
  MatBatchArgs a = batch_args(ctx);
  a.n = ctx->ne * ctx->nq;
  a.hq = hq;
  a.eperm = get<int32_t>(ctx, "eperm");
  a.st_old = get<float>(ctx, "state_old");
  a.st_new = get<float>(ctx, "state_new");
  a.qpmat = qpmat;
  a.flag = get<int>(ctx, "flag");
  a.mp.phys = ctx->mp.phys == FECX_J2_SURROGATE ? FECX_J2 : FECX_NEOHOOKEAN;
  a.accept = (uint8_t*)fecx_api::get<uint8_t>(ctx, "sg_accept");
  if (!a.accept || fecx_api::get_bytes(ctx, "sg_accept") < (size_t)a.n) {
    a.accept = (uint8_t*)dev(ctx, "sg_accept", (size_t)a.n);
    ctx->sg_frozen_valid = false;
  }
  a.freeze = ctx->sg_freeze && ctx->sg_frozen_valid ? 1 : 0;
  ctx->sg_frozen_valid = true;
  HIPCHK(hipMemsetAsync(a.nfallback, 0, sizeof(unsigned long long), s));
  mlp_launch_material(ctx->sg_arch, true, a, s);
  ctx->sg_points = a.n;
}

void surrogate_qp_finish(fecx_ctx* ctx) {
  unsigned long long nfb = 0;
  HIPCHK(hipMemcpy(&nfb, get<unsigned long long>(ctx, "sg_nfb"), sizeof(nfb), hipMemcpyDeviceToHost));
  ctx->sg_fallback = (int64_t)nfb;
}

}  // namespace fecx

extern "C" {

int fecx_surrogate_init(fecx_ctx* ctx, int arch, uint64_t seed) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(arch == FECX_MLP_J2 || arch == FECX_MLP_NH, FECX_ERR_ARG, "unknown surrogate architecture");
    const fecx::ArchInfo ai = fecx::arch_info(arch);
    ctx->sg_arch = arch;
    ctx->sg_step = 0;
    // He-normal weights, zero biases
    std::vector<float> w(ai.nparams, 0.0f);
    SplitMix rng{seed};
    int64_t o = 0;
    for (int l = 0; l < ai.nl; ++l) {
      const double sd = std::sqrt(2.0 / ai.d[l]);
      for (int64_t k = 0; k < (int64_t)ai.d[l + 1] * ai.d[l]; ++k) w[o + k] = (float)(sd * rng.normal());
      o += (int64_t)ai.d[l + 1] * ai.d[l] + ai.d[l + 1];
    }
    upload(ctx, "sg_w", w);
    const size_t pb = sizeof(float) * ai.nparams;
    HIPCHK(hipMemset(dev(ctx, "sg_m1", pb), 0, pb));
    HIPCHK(hipMemset(dev(ctx, "sg_m2", pb), 0, pb));
    const size_t wfb = 16 * ai.wf_vecs, wbb = 16 * ai.wb_vecs, bb = sizeof(float) * ai.nbias;
    HIPCHK(hipMemset(dev(ctx, "sg_wf", wfb), 0, wfb));
    HIPCHK(hipMemset(dev(ctx, "sg_wb", wbb), 0, wbb));
    HIPCHK(hipMemset(dev(ctx, "sg_bias", bb), 0, bb));
    upload(ctx, "sg_in_shift", std::vector<float>(ai.nfeat, 0.0f));
    upload(ctx, "sg_in_scale", std::vector<float>(ai.nfeat, 1.0f));
    upload(ctx, "sg_out_shift", std::vector<float>(ai.out_w, 0.0f));
    upload(ctx, "sg_out_scale", std::vector<float>(ai.out_w, 1.0f));
    repack(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_surrogate_num_params(fecx_ctx* ctx, int64_t* n) {
  return guard(ctx, [&] {
    require(ctx->sg_arch >= 0, FECX_ERR_STATE, "fecx_surrogate_init has not been called");
    require(n != nullptr, FECX_ERR_ARG, "null output");
    *n = fecx::arch_info(ctx->sg_arch).nparams;
  });
}

int fecx_surrogate_get_params(fecx_ctx* ctx, float* params) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(params != nullptr, FECX_ERR_ARG, "null output");
    const int64_t np = fecx::arch_info(ctx->sg_arch).nparams;
    HIPCHK(hipMemcpy(params, get<float>(ctx, "sg_w"), sizeof(float) * np, hipMemcpyDeviceToHost));
  });
}

int fecx_surrogate_set_params(fecx_ctx* ctx, const float* params) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(params != nullptr, FECX_ERR_ARG, "null input");
    const int64_t np = fecx::arch_info(ctx->sg_arch).nparams;
    HIPCHK(hipMemcpy(get<float>(ctx, "sg_w"), params, sizeof(float) * np, hipMemcpyHostToDevice));
    repack(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_surrogate_set_scaling(fecx_ctx* ctx, const float* in_shift, const float* in_scale, const float* out_shift,
                               const float* out_scale) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(in_shift && in_scale && out_shift && out_scale, FECX_ERR_ARG, "null scaling array");
    const fecx::ArchInfo ai = fecx::arch_info(ctx->sg_arch);
    for (int i = 0; i < ai.nfeat; ++i) require(in_scale[i] > 0.0f, FECX_ERR_ARG, "scales must be positive");
    for (int i = 0; i < ai.out_w; ++i) require(out_scale[i] > 0.0f, FECX_ERR_ARG, "scales must be positive");
    upload(ctx, "sg_in_shift", std::vector<float>(in_shift, in_shift + ai.nfeat));
    upload(ctx, "sg_in_scale", std::vector<float>(in_scale, in_scale + ai.nfeat));
    upload(ctx, "sg_out_shift", std::vector<float>(out_shift, out_shift + ai.out_w));
    upload(ctx, "sg_out_scale", std::vector<float>(out_scale, out_scale + ai.out_w));
  });
}

int fecx_surrogate_get_scaling(fecx_ctx* ctx, float* in_shift, float* in_scale, float* out_shift, float* out_scale) {
  return guard(ctx, [&] {
    require_net(ctx);
    const fecx::ArchInfo ai = fecx::arch_info(ctx->sg_arch);
    auto cp = [&](float* dst, const char* name, int n) {
      if (dst) HIPCHK(hipMemcpy(dst, get<float>(ctx, name), sizeof(float) * n, hipMemcpyDeviceToHost));
    };
    cp(in_shift, "sg_in_shift", ai.nfeat);
    cp(in_scale, "sg_in_scale", ai.nfeat);
    cp(out_shift, "sg_out_shift", ai.out_w);
    cp(out_scale, "sg_out_scale", ai.out_w);
  });
}

int fecx_surrogate_fit_scaling(fecx_ctx* ctx, const float* X, const float* T, int64_t n) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(X && T && n > 0, FECX_ERR_ARG, "empty sample set");
    const fecx::ArchInfo ai = fecx::arch_info(ctx->sg_arch);
    const float* feat = X;
    int fw = ai.in_w;
    if (ctx->sg_arch == FECX_MLP_J2) {
      float* f = (float*)dev_atleast(ctx, "sg_feat", sizeof(float) * n * ai.nfeat);
      fecx::mlp_launch_j2_features(X, n, f, ctx->stream);
      feat = f;
      fw = ai.nfeat;
    }
    fecx::mlp_launch_column_stats(feat, n, fw, ai.nfeat, get<float>(ctx, "sg_in_shift"), get<float>(ctx, "sg_in_scale"),
                                  ctx->stream);
    fecx::mlp_launch_column_stats(T, n, ai.out_w, ai.out_w, get<float>(ctx, "sg_out_shift"),
                                  get<float>(ctx, "sg_out_scale"), ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_surrogate_forward(fecx_ctx* ctx, const float* X, int64_t n, float* Y) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(n >= 0 && (n == 0 || (X && Y)), FECX_ERR_ARG, "bad arrays");
    fecx::MatBatchArgs a = batch_args(ctx);
    a.n = n;
    a.in = X;
    a.out = Y;
    fecx::mlp_launch_forward(ctx->sg_arch, a, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_surrogate_gradient(fecx_ctx* ctx, const float* X, const float* T, int64_t n, double* loss, float* grad) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(X && T && n > 0 && grad, FECX_ERR_ARG, "bad arrays");
    require(n <= (int64_t(1) << 30), FECX_ERR_ARG, "too many samples for one pass");
    double* dl = (double*)dev_atleast(ctx, "sg_loss", sizeof(double));
    grad_pass(ctx, X, T, nullptr, n, grad, false, 0.0f, dl);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (loss) HIPCHK(hipMemcpy(loss, dl, sizeof(double), hipMemcpyDeviceToHost));
  });
}

int fecx_surrogate_train(fecx_ctx* ctx, const float* X, const float* T, int64_t n, int epochs, int batch, float lr,
                         uint64_t seed, double* loss_per_epoch) {
  return guard(ctx, [&] {
    require_net(ctx);
    require(X && T && n > 0 && epochs >= 0 && batch >= 32 && lr > 0.0f, FECX_ERR_ARG, "bad training arguments");
    require(n < (int64_t(1) << 31), FECX_ERR_ARG, "at most 2^31 - 1 samples");
    const int64_t nb = (n + batch - 1) / batch;
    std::vector<int32_t> perm(n);
    int32_t* dperm = (int32_t*)dev_atleast(ctx, "sg_perm", sizeof(int32_t) * n);
    double* dl = (double*)dev_atleast(ctx, "sg_epoch_loss", sizeof(double) * nb);
    std::vector<double> hl(nb);
    SplitMix rng{seed};
    for (int ep = 0; ep < epochs; ++ep) {
      std::iota(perm.begin(), perm.end(), 0);
      for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int64_t)(rng.next() % (uint64_t)(i + 1))]);
      HIPCHK(hipMemcpyAsync(dperm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
      for (int64_t b = 0; b < nb; ++b) {
        const int64_t nbatch = std::min<int64_t>(batch, n - b * batch);
        grad_pass(ctx, X, T, dperm + b * batch, nbatch, nullptr, true, lr, dl + b);
      }
      HIPCHK(hipStreamSynchronize(ctx->stream));  // the permutation buffer is reused by the next epoch
      HIPCHK(hipMemcpy(hl.data(), dl, sizeof(double) * nb, hipMemcpyDeviceToHost));
      if (loss_per_epoch) {
        double s = 0.0;
        for (double v : hl) s += v;
        loss_per_epoch[ep] = s / (double)nb;
      }
    }
  });
}

int fecx_surrogate_set_tolerance(fecx_ctx* ctx, float tol) {
  return guard(ctx, [&] {
    require(tol > 0.0f, FECX_ERR_ARG, "tolerance must be positive");
    ctx->sg_tol = tol;
  });
}

int fecx_surrogate_freeze(fecx_ctx* ctx, int on) {
  return guard(ctx, [&] { ctx->sg_freeze = on != 0; });
}

int fecx_surrogate_enable(fecx_ctx* ctx, int on) {
  return guard(ctx, [&] { ctx->sg_enabled = on != 0; });
}

int fecx_surrogate_fallbacks(fecx_ctx* ctx, int64_t* n_fallback, int64_t* n_points) {
  return guard(ctx, [&] {
    if (n_fallback) *n_fallback = ctx->sg_fallback;
    if (n_points) *n_points = ctx->sg_points;
  });
}

int fecx_material_update(fecx_ctx* ctx, int kind, const float* in, int64_t n, float* out) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(ctx->has_material, FECX_ERR_STATE, "fecx_set_material has not been called");
    require(n >= 0 && (n == 0 || (in && out)), FECX_ERR_ARG, "bad arrays");
    const bool j2 = kind == FECX_J2 || kind == FECX_J2_SURROGATE;
    const bool nh = kind == FECX_NEOHOOKEAN || kind == FECX_NH_SURROGATE;
    require(j2 || nh, FECX_ERR_UNSUPPORTED, "batched update exists for J2 and neo-Hookean only");
    const int ph = ctx->mp.phys;
    require(j2 ? (ph == FECX_J2 || ph == FECX_J2_SURROGATE) : (ph == FECX_NEOHOOKEAN || ph == FECX_NH_SURROGATE),
            FECX_ERR_STATE, "the context's material is not of this family");
    const bool surr = kind == FECX_J2_SURROGATE || kind == FECX_NH_SURROGATE;
    const int arch = j2 ? FECX_MLP_J2 : FECX_MLP_NH;
    if (surr) require(ctx->sg_arch == arch, FECX_ERR_STATE, "the surrogate network for this material is not initialised");
    fecx::MatBatchArgs a{};
    if (ctx->sg_arch >= 0) a = batch_args(ctx);
    a.n = n;
    a.in = in;
    a.out = out;
    a.mp = ctx->mp;
    a.mp.phys = j2 ? FECX_J2 : FECX_NEOHOOKEAN;
    a.tol = ctx->sg_tol;
    if (surr) HIPCHK(hipMemsetAsync(a.nfallback, 0, sizeof(unsigned long long), ctx->stream));
    else a.nfallback = nullptr;
    fecx::mlp_launch_material(arch, surr, a, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (surr) {
      unsigned long long nfb = 0;
      HIPCHK(hipMemcpy(&nfb, a.nfallback, sizeof(nfb), hipMemcpyDeviceToHost));
      ctx->sg_fallback = (int64_t)nfb;
      ctx->sg_points = n;
    }
  });
}

int fecx_qp_gradients(fecx_ctx* ctx, const double* U, float* Hq) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(ctx->has_material && ctx->has_pattern, FECX_ERR_STATE, "material and pattern must be set");
    require(U && Hq, FECX_ERR_ARG, "null arrays");
    require(ctx->nf == 3 && ctx->nd == 3, FECX_ERR_UNSUPPORTED, "quadrature-point gradients for 3D vector fields only");
    const int64_t nqp = ctx->ne * ctx->nq;
    fecx::ElementArgs ea{};
    ea.ebeg = 0;
    ea.ne = ctx->ne;
    ea.conn = get<int32_t>(ctx, "conn");
    ea.eperm = get<int32_t>(ctx, "eperm");
    ea.xhi = get<float>(ctx, "xhi");
    ea.xlo = get<float>(ctx, "xlo");
    ea.U = U;
    ea.flag = get<int>(ctx, "flag");
    ea.mp = ctx->mp;
    ea.fp64_geometry = ctx->fp64_geometry ? 1 : 0;
    ea.hq = (float*)dev_atleast(ctx, "hq", sizeof(float) * nqp * 9);
    HIPCHK(hipMemsetAsync(ea.flag, 0, sizeof(int), ctx->stream));
    ctx->flag_dirty = true;  // not read back here: the next assembly clears it first
    require(fecx::launch_element(ctx->kind, ctx->nf, fecx::MODE_QPGRAD, ea, ctx->stream) == 0, FECX_ERR_UNSUPPORTED,
            "no gradient kernel for this element kind");
    // storage order -> element order: row e of Hq is row epos[e] of the storage-ordered buffer
    fecx::mlp_launch_gather(ctx->ne, ctx->nq * 9, get<int32_t>(ctx, "epos"), ea.hq, Hq, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int fecx_j2_generate_samples(fecx_ctx* ctx, int64_t npaths, int steps, float max_norm, uint64_t seed, float* X,
                             float* T) {
  return guard(ctx, [&] {
    check_device_ctx(ctx);
    require(ctx->has_material && (ctx->mp.phys == FECX_J2 || ctx->mp.phys == FECX_J2_SURROGATE), FECX_ERR_STATE,
            "the context's material is not J2");
    require(npaths >= 0 && steps > 0 && max_norm > 0.0f && (npaths == 0 || (X && T)), FECX_ERR_ARG, "bad arguments");
    fecx::mlp_launch_j2_paths(npaths, steps, max_norm, seed, ctx->mp, X, T, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
