// C-ABI entry points of libfecx.so (declared in include/fecx.h, each with its reference citation).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "api_util.h"
#include "surrogate.h"

namespace {
using namespace fecx_api;

// kernel profiling with HIP events on the context's stream
struct Prof {
  fecx_ctx* ctx;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev;
  explicit Prof(fecx_ctx* c) : ctx(c) {}
  template <class F>
  void run(const char* name, hipStream_t st, F f) {
    if (!ctx || !ctx->profiling) { f(); return; }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, st));
    f();
    HIPCHK(hipEventRecord(b, st));
    ev.push_back({name, {a, b}});
  }
  void finish() {
    if (!ctx) return;
    for (auto& x : ev) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, x.second.first, x.second.second));
      auto& pe = ctx->prof[x.first];
      if (pe.launches == 0 && std::find(ctx->prof_names.begin(), ctx->prof_names.end(), x.first) == ctx->prof_names.end())
        ctx->prof_names.push_back(x.first);
      pe.launches++;
      pe.ms += ms;
      (void)hipEventDestroy(x.second.first);
      (void)hipEventDestroy(x.second.second);
    }
    ev.clear();
  }
};

__global__ void zero_int_kernel(int* p) { *p = 0; }

void check_device(fecx_ctx* ctx) {
  require(ctx->device >= 0, FECX_ERR_STATE, "host-only context (device -1) cannot run kernels");
}

void check_ready(fecx_ctx* ctx) {
  check_device(ctx);
  require(ctx->has_mesh, FECX_ERR_STATE, "fecx_set_mesh has not been called");
  require(ctx->has_dofs, FECX_ERR_STATE, "fecx_set_dofs has not been called");
  require(ctx->has_pattern, FECX_ERR_STATE, "fecx_build_pattern has not been called");
  require(ctx->has_material, FECX_ERR_STATE, "fecx_set_material has not been called");
}

fecx::ElementArgs element_args(fecx_ctx* ctx, const double* U, const double* V) {
  fecx::ElementArgs a{};
  a.ebeg = 0;
  a.ne = ctx->ne;
  a.conn = get<int32_t>(ctx, "conn");
  a.eperm = get<int32_t>(ctx, "eperm");
  a.xhi = get<float>(ctx, "xhi");
  a.xlo = get<float>(ctx, "xlo");
  a.U = U;
  a.V = V;
  a.st_old = get<float>(ctx, "state_old");
  a.st_new = get<float>(ctx, "state_new");
  a.Rel = get<float>(ctx, "Rel");
  a.Kel = get<float>(ctx, "Kel");
  a.flag = get<int>(ctx, "flag");
  a.mp = ctx->mp;
  a.fp64_geometry = ctx->fp64_geometry ? 1 : 0;
  return a;
}

fecx::ReduceArgs reduce_args(fecx_ctx* ctx) {
  fecx::ReduceArgs r{};
  r.p0 = 0;
  r.p1 = ctx->nn;
  r.rowinfo = get<int4>(ctx, "rowinfo");
  r.nnpe = ctx->nnpe;
  r.maxdeg = ctx->hp.maxdeg;
  r.nlight = ctx->hp.nlight;
  r.npad = ctx->nnpe == 8 && ctx->nf == 1 ? ctx->hp.nlight : 0;  // padded light rows (see fecx_build_pattern)
  r.maxdeg_light = ctx->hp.maxdeg_light;
  r.vent = get<int32_t>(ctx, "vent");
  r.vslot = get<uint8_t>(ctx, "vslot");
  r.kel_off = get<int16_t>(ctx, "kel_off");
  r.merged = ctx->merged_rows ? get<int2>(ctx, "rmerge") : nullptr;
  r.nmerged = (int64_t)(get_bytes(ctx, "rmerge") / sizeof(int2));
  r.Kel = get<float>(ctx, "Kel");
  r.Rel = get<float>(ctx, "Rel");
  return r;
}

// Number of pipeline chunks of a matrix assembly. One by default: on MI355X the overlapped pieces measured slower
// than back-to-back launches (hex8 neo-Hookean 96^3: 2.04 ms one piece, 2.15-2.38 ms with 4-32 pieces), the two
// kernels competing for the same CU resources.
int pipeline_chunks(fecx_ctx* ctx, int mode) {
  if (!(mode & fecx::MODE_K)) return 1;
  if (ctx->hp.nlight < ctx->nn) return 1;  // two row buckets: the chunk keys assume one Morton-ordered bucket
  return ctx->chunks > 0 ? ctx->chunks : 1;
}

// A hex8 mesh numbered as StructuredMesh numbers it (src/meshes/StructuredMesh.jl:85-131): node (i, j, k) is
// i + nx (j + ny k) (0-based), element (ex, ey, ez) is ez + nez (ey + ney ex) with local nodes in the order of
// StructuredMesh.jl:116-123. Finds the node counts from the first element and checks every element against them.
bool detect_structured_hex8(fecx_ctx* ctx) {
  ctx->structured = false;
  if (ctx->kind != FECX_HEX8 || ctx->block_start.size() != 2 || ctx->ne < 1) return false;
  const int32_t* c = ctx->conn0.data();
  if (c[0] != 0 || c[1] != 1) return false;
  const int64_t nx = c[3], nxy = c[4];
  if (nx < 2 || nxy < 2 * nx || nxy % nx != 0 || ctx->nn % nxy != 0) return false;
  const int64_t ny = nxy / nx, nz = ctx->nn / nxy;
  if (nz < 2 || (nx - 1) * (ny - 1) * (nz - 1) != ctx->ne || nx >= (1 << 20) || ny >= (1 << 20)) return false;
  const int64_t nex = nx - 1, ney = ny - 1, nez = nz - 1;
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int64_t e = 0; e < ctx->ne; ++e) {
    const int64_t ez = e % nez, ey = (e / nez) % ney, ex = e / (nez * ney);
    for (int a = 0; a < 8; ++a) {
      const int64_t n = (ex + off[a][0]) + nx * ((ey + off[a][1]) + ny * (ez + off[a][2]));
      if (c[e * 8 + a] != n) return false;
    }
  }
  ctx->snx = (int)nx; ctx->sny = (int)ny; ctx->snz = (int)nz;
  ctx->structured = true;
  return true;
}

// Per node {first CSR block of its row, mask of the 27 neighbours (dx, dy, dz) present in the row}, bit
// (dz + 1) 9 + (dy + 1) 3 + (dx + 1): the neighbour exists and both nodes are free. Checked against the pattern's
// row degrees; returns false (and the fused path stays off) on a mismatch.
bool build_nodeinfo(fecx_ctx* ctx) {
  const auto& hp = ctx->hp;
  const int64_t nx = ctx->snx, ny = ctx->sny, nz = ctx->snz;
  std::vector<int32_t> info((size_t)ctx->nn * 2, 0);
  for (int64_t n = 0; n < ctx->nn; ++n) {
    const int32_t r = hp.free_rank[n];
    if (r < 0) continue;
    const int64_t i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
    uint32_t mask = 0;
    for (int d = 0; d < 27; ++d) {
      const int64_t mi = i + d % 3 - 1, mj = j + (d / 3) % 3 - 1, mk = k + d / 9 - 1;
      if (mi < 0 || mi >= nx || mj < 0 || mj >= ny || mk < 0 || mk >= nz) continue;
      if (hp.free_rank[mi + nx * (mj + ny * mk)] >= 0) mask |= 1u << d;
    }
    if ((int64_t)__builtin_popcount(mask) != hp.bptr[r + 1] - hp.bptr[r]) return false;
    if (hp.bptr[r] >= (int64_t(1) << 31)) return false;
    info[2 * n] = (int32_t)hp.bptr[r];
    info[2 * n + 1] = (int32_t)mask;
  }
  upload(ctx, "nodeinfo", info);
  return true;
}

// The scatter design a matrix assembly takes (FECX_SCATTER_*): the fused kernel for residual + Jacobian (or Jacobian)
// assemblies of a one-block structured hex8 mesh with a plain material when asked for, or by default for the
// vector-valued materials (on the scalar Poisson problem the two-phase path is faster: its element work is cheap and
// the fused kernel's halo recompute and one-work-group-per-CU marching do not pay); the atomic and coloured designs
// only where they are built (structured hex8, Poisson or neo-Hookean) and never for the allocation-free entry point
// (their zero fill is a memset); the two-phase path everywhere else.
int scatter_path(fecx_ctx* ctx, int mode, bool surrogate, bool noalloc) {
  if (!ctx->structured || surrogate || ctx->block_mp.size() != 1) return FECX_SCATTER_TWO_PHASE;
  if (mode != (fecx::MODE_R | fecx::MODE_K) && mode != fecx::MODE_K) return FECX_SCATTER_TWO_PHASE;
  const int ph = ctx->mp.phys;
  const bool fused_ok = (ctx->nf == 1 && ph == FECX_POISSON) ||
                        (ctx->nf == 3 && (ph == FECX_LINEAR_ELASTIC || ph == FECX_NEOHOOKEAN || ph == FECX_J2));
  const bool direct_ok = (ctx->nf == 1 && ph == FECX_POISSON) || (ctx->nf == 3 && ph == FECX_NEOHOOKEAN);
  if (ctx->scatter == FECX_SCATTER_FUSED && fused_ok) return FECX_SCATTER_FUSED;
  if (ctx->scatter == FECX_SCATTER_AUTO && fused_ok && ctx->nf == 3) return FECX_SCATTER_FUSED;
  if ((ctx->scatter == FECX_SCATTER_ATOMIC || ctx->scatter == FECX_SCATTER_COLOURED) && direct_ok && !noalloc)
    return ctx->scatter;
  return FECX_SCATTER_TWO_PHASE;
}

// With a partial Dirichlet set: the dof-level CSR values out of the node-block values the kernels assembled
// (stream-ordered after them).
void compact_values(fecx_ctx* ctx, const double* full, double* out) {
  if (!ctx->partial || !out || ctx->nnz == 0) return;
  fecx::launch_gather_unknowns(ctx->nnz, get<int64_t>(ctx, "cmap"), full, out, ctx->stream);
  HIPCHK(hipGetLastError());
}

// The numeric flag before a launch: cleared on the device only if the last read-back did not find it clear.
int* arm_flag(fecx_ctx* ctx) {
  int* flag = get<int>(ctx, "flag");
  if (ctx->flag_dirty) zero_int_kernel<<<1, 1, 0, ctx->stream>>>(flag);
  ctx->flag_dirty = true;  // until this call reads it back as clear
  require(ctx->flag_host != nullptr, FECX_ERR_STATE, "no host flag word");
  return flag;
}

// Reads the flag back (pinned word, stream-ordered after the kernels) and waits for the stream.
int read_flag(fecx_ctx* ctx, const int* flag) {
  HIPCHK(hipMemcpyAsync(ctx->flag_host, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int h = *ctx->flag_host;
  ctx->flag_dirty = h != 0;
  return h;
}

// Runs the element kernel for `mode`, then the requested reductions; syncs and checks the flag.
// With more than one chunk, the Morton-ordered row nodes are split into `k` ranges of positions [p_j, p_j+1); element
// chunk j is the storage range of the elements whose smallest node rank lies in that range (contiguous, elements are
// stored sorted by that key). A node only meets elements whose key is <= its own rank, so reduction chunk j needs
// element chunks 0..j: element chunks run in order on stream2, reduction chunk j on the context's stream after an
// event recorded behind element chunk j. Element chunk j + 1 (arithmetic-bound) overlaps reduction chunk j
// (memory-bound), which reads element storage written moments earlier.
// With `noalloc`, nothing is allocated: no profiling events, one piece, no surrogate pass; the two-phase path keeps its
// element-private outputs in the per-device scratch array (noalloc.hip) instead of the context's buffers, and the
// fused kernel (structured hex8 meshes) needs none, so no scratch size limits it.
void assemble(fecx_ctx* ctx, int mode, const double* U, const double* V, double* R, double* vals,
              bool noalloc = false) {
  check_ready(ctx);
  require(U != nullptr, FECX_ERR_ARG, "U is NULL");
  double* const vals_out = vals;  // the caller's CSR; with a partial Dirichlet set the kernels fill vals_full
  if (ctx->partial && vals) vals = get<double>(ctx, "vals_full");
  const bool surrogate_phys = ctx->mp.phys == FECX_J2_SURROGATE || ctx->mp.phys == FECX_NH_SURROGATE;
  const bool surrogate = surrogate_phys && ctx->sg_enabled;
  float *kel = get<float>(ctx, "Kel"), *rel = get<float>(ctx, "Rel");
  const int path = scatter_path(ctx, mode, surrogate_phys, noalloc);
  if (noalloc) require(!surrogate, FECX_ERR_UNSUPPORTED, "the allocation-free entry point does not run surrogate materials");
  if (noalloc && path == FECX_SCATTER_TWO_PHASE) {  // the fused kernel keeps no element-private outputs
    const size_t kb = sizeof(float) * (size_t)ctx->ne * ctx->pairs.npairs * ctx->nf * ctx->nf;
    const size_t rb = sizeof(float) * (size_t)ctx->ne * ctx->nnpe * ctx->nf;
    const size_t kb_al = (kb + 255) / 256 * 256;
    require(kb_al + rb <= fecx::NOALLOC_SCRATCH_BYTES, FECX_ERR_ARG,
            "the allocation-free entry point needs " + std::to_string(kb_al + rb) + " bytes of element scratch; it "
            "holds at most " + std::to_string(fecx::NOALLOC_SCRATCH_BYTES) + " (use fecx_assemble_vector_matrix)");
    unsigned char* base = (unsigned char*)fecx::noalloc_scratch(ctx->device);
    require(base != nullptr, FECX_ERR_STATE, "no scratch address for this device");
    kel = (float*)base;
    rel = (float*)(base + kb_al);
  }
  Prof prof(ctx);
  if (noalloc) prof.ctx = nullptr;  // no HIP events
  int* flag = arm_flag(ctx);
  fecx::ElementArgs ea = element_args(ctx, U, V);
  ea.Kel = kel;
  ea.Rel = rel;
  if (surrogate_phys && !surrogate) ea.mp.phys = ctx->mp.phys == FECX_J2_SURROGATE ? FECX_J2 : FECX_NEOHOOKEAN;
  if (ctx->ns > 0) require(ea.st_old != nullptr, FECX_ERR_STATE, "state not allocated");
  if (surrogate) {
    // quadrature-point gradients, then the batched surrogate material pass (network, check, plain fallback)
    require(ctx->sg_arch == (ctx->mp.phys == FECX_J2_SURROGATE ? FECX_MLP_J2 : FECX_MLP_NH), FECX_ERR_STATE,
            "the surrogate network for this material has not been initialised");
    const int64_t nqp = ctx->ne * ctx->nq;
    const int nc = ctx->mp.phys == FECX_J2_SURROGATE ? 9 : 10;
    ea.hq = (float*)dev_atleast(ctx, "hq", sizeof(float) * nqp * 9);
    ea.qpmat = (float*)dev_atleast(ctx, "qpmat", sizeof(float) * nqp * (9 + nc));
    int rc = 0;
    prof.run("qp_gradients", ctx->stream, [&] { rc = fecx::launch_element(ctx->kind, ctx->nf, fecx::MODE_QPGRAD, ea, ctx->stream); });
    require(rc == 0, FECX_ERR_UNSUPPORTED, "no gradient kernel for this element kind / nf");
    prof.run("surrogate_material", ctx->stream, [&] { fecx::surrogate_qp_pass(ctx, ea.hq, (float*)ea.qpmat, ctx->stream); });
    HIPCHK(hipGetLastError());
  }
  if (path != FECX_SCATTER_TWO_PHASE) {
    fecx::FusedArgs fa{};
    fa.nx = ctx->snx; fa.ny = ctx->sny; fa.nz = ctx->snz;
    fa.epos = get<int32_t>(ctx, "epos");
    fa.nodeinfo = get<int2>(ctx, "nodeinfo");
    fa.vals = (mode & fecx::MODE_K) ? vals : nullptr;
    fa.R = (mode & fecx::MODE_R) ? R : nullptr;
    int rc = 0;
    if (path == FECX_SCATTER_FUSED) {
      prof.run("fused", ctx->stream, [&] { rc = fecx::launch_fused_hex8(ctx->nf, ea, fa, ctx->stream); });
    } else {  // atomic or coloured: zero-fill, then every element adds its blocks and rows in place
      const size_t rbytes = sizeof(double) * (size_t)ctx->nn * ctx->nf;
      prof.run(path == FECX_SCATTER_ATOMIC ? "atomic" : "coloured", ctx->stream, [&] {
        if (fa.vals) HIPCHK(hipMemsetAsync(fa.vals, 0, sizeof(double) * (size_t)ctx->nnz_full, ctx->stream));
        if (fa.R) HIPCHK(hipMemsetAsync(fa.R, 0, rbytes, ctx->stream));
        rc = fecx::launch_scatter_hex8(ctx->nf, path == FECX_SCATTER_ATOMIC ? 1 : 2, ea, fa, ctx->stream);
      });
    }
    require(rc == 0, FECX_ERR_UNSUPPORTED, "no kernel of this scatter design for this nf / physics");
    HIPCHK(hipGetLastError());
    compact_values(ctx, vals, vals_out);
    const int hflag = read_flag(ctx, flag);
    prof.finish();
    require(hflag == 0, FECX_ERR_NUMERIC, "element kernel met a non-positive Jacobian determinant or a NaN");
    return;
  }
  fecx::ReduceArgs ra = reduce_args(ctx);
  ra.Kel = kel;
  ra.Rel = rel;
  if (mode & (fecx::MODE_R | fecx::MODE_ACTION_MF | fecx::MODE_ACTION_ASM | fecx::MODE_DIAG)) ra.R = R;
  if (mode & fecx::MODE_K) ra.vals = vals;
  const int nblocks = (int)ctx->block_mp.size();
  const int k = (nblocks > 1 || noalloc) ? 1 : pipeline_chunks(ctx, mode);
  hipStream_t es = ctx->stream;
  if (k > 1) {
    if (!ctx->stream2) HIPCHK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    while ((int)ctx->events.size() < k + 1) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ctx->events.push_back(e);
    }
    es = ctx->stream2;
    HIPCHK(hipEventRecord(ctx->events[k], ctx->stream));  // element storage is free once earlier work is done
    HIPCHK(hipStreamWaitEvent(es, ctx->events[k], 0));
  }
  for (int j = 0; j < k; ++j) {
    const int64_t p0 = ctx->nn * j / k, p1 = ctx->nn * (j + 1) / k;
    int rc = 0;
    if (nblocks == 1) {
      ea.ebeg = ctx->hp.ekey[p0];
      ea.ne = ctx->hp.ekey[p1];
      prof.run("element", es, [&] { rc = fecx::launch_element(ctx->kind, ctx->nf, mode, ea, es); });
    } else {  // foreach_block (src/Utils.jl:171-189): each block's storage range with its own material
      for (int b = 0; b < nblocks && rc == 0; ++b) {
        fecx::ElementArgs eb = ea;
        eb.ebeg = ctx->hp.block_pos[b];
        eb.ne = ctx->hp.block_pos[b + 1];
        eb.mp = ctx->block_mp[b];
        prof.run("element", es, [&] { rc = fecx::launch_element(ctx->kind, ctx->nf, mode, eb, es); });
      }
    }
    require(rc == 0, FECX_ERR_UNSUPPORTED, "no element kernel for this element kind / nf / physics");
    HIPCHK(hipGetLastError());
    if (k > 1) {
      HIPCHK(hipEventRecord(ctx->events[j], es));
      HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->events[j], 0));
    }
    ra.p0 = p0;
    ra.p1 = p1;
    if (k > 1) ra.merged = nullptr;  // the merged launch's chunk table covers all rows
    prof.run("reduce", ctx->stream, [&] { rc = fecx::launch_reduce(ctx->nf, ra, ctx->stream); });
    require(rc == 0, FECX_ERR_UNSUPPORTED, "no reduction kernel for this element kind / nf");
    HIPCHK(hipGetLastError());
  }
  compact_values(ctx, vals, vals_out);
  const int hflag = read_flag(ctx, flag);
  prof.finish();
  if (surrogate) fecx::surrogate_qp_finish(ctx);
  require(hflag == 0, FECX_ERR_NUMERIC, "element kernel met a non-positive Jacobian determinant or a NaN");
}

// Mass-type assembly: the mass kernel for `what`, then the reduction into the CSR values (MASS_MATRIX) or the full
// field R (the others). Needs the mesh, the DOFs and the pattern, not a material.
void assemble_mass(fecx_ctx* ctx, int what, double density, const double* V, double* R, double* vals) {
  check_device(ctx);
  require(ctx->has_mesh && ctx->has_dofs && ctx->has_pattern, FECX_ERR_STATE,
          "mesh, dofs and pattern must be set before a mass assembly");
  require(std::isfinite(density), FECX_ERR_ARG, "density must be finite");
  Prof prof(ctx);
  int* flag = arm_flag(ctx);
  fecx::ElementArgs ea = element_args(ctx, nullptr, V);
  int rc = 0;
  prof.run("mass_element", ctx->stream, [&] { rc = fecx::launch_mass(ctx->kind, ctx->nf, what, ea, (float)density, ctx->stream); });
  require(rc == 0, FECX_ERR_UNSUPPORTED, "no mass kernel for this element kind / nf");
  HIPCHK(hipGetLastError());
  fecx::ReduceArgs ra = reduce_args(ctx);
  ra.R = R;
  ra.vals = (ctx->partial && vals) ? get<double>(ctx, "vals_full") : vals;
  prof.run("reduce", ctx->stream, [&] { rc = fecx::launch_reduce(ctx->nf, ra, ctx->stream); });
  require(rc == 0, FECX_ERR_UNSUPPORTED, "no reduction kernel for this element kind / nf");
  HIPCHK(hipGetLastError());
  compact_values(ctx, ra.vals, vals);
  const int hflag = read_flag(ctx, flag);
  prof.finish();
  require(hflag == 0, FECX_ERR_NUMERIC, "mass kernel met a non-positive Jacobian determinant");
}

}  // namespace

extern "C" {

int fecx_create(fecx_ctx** out, int device) {
  if (!out) return FECX_ERR_ARG;
  *out = nullptr;
  if (device < -1) return FECX_ERR_ARG;
  if (device >= 0) {  // device -1: host-only context (pattern and index maps only)
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device >= n) return FECX_ERR_HIP;
  }
  fecx_ctx* ctx = new fecx_ctx();
  ctx->device = device;
  int rc = guard(ctx, [&] {
    (void)dev(ctx, "flag", sizeof(int));
    if (device >= 0) {
      HIPCHK(hipHostMalloc((void**)&ctx->flag_host, sizeof(int), hipHostMallocDefault));
      // the allocation-free entry point's per-device scratch address, looked up once per device, device current
      require(fecx::noalloc_setup(device) == FECX_OK, FECX_ERR_HIP, "scratch symbol lookup failed");
    }
  });
  if (rc) { delete ctx; return rc; }
  *out = ctx;
  return FECX_OK;
}

int fecx_destroy(fecx_ctx* ctx) {
  if (!ctx) return FECX_OK;
  if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
  for (auto& kv : ctx->buf)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->flag_host) (void)hipHostFree(ctx->flag_host);
  delete ctx;
  return FECX_OK;
}

const char* fecx_last_error(const fecx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fecx_set_stream(fecx_ctx* ctx, void* s) {
  return guard(ctx, [&] { ctx->stream = (hipStream_t)s; });
}

int fecx_set_mesh(fecx_ctx* ctx, int elem_kind, int nd, int64_t nn, const double* coords, int64_t ne,
                  const int64_t* conn) {
  return fecx_set_mesh_blocks(ctx, elem_kind, nd, nn, coords, 1, &ne, conn);
}

int fecx_set_mesh_blocks(fecx_ctx* ctx, int elem_kind, int nd, int64_t nn, const double* coords, int nblocks,
                         const int64_t* block_ne, const int64_t* conn) {
  return guard(ctx, [&] {
    fecx::ElemInfo ei = fecx::elem_info(elem_kind);
    require(ei.kind >= 0, FECX_ERR_UNSUPPORTED, "unsupported element kind");
    require(nd == ei.nd, FECX_ERR_ARG, "coordinate dimension does not match the element kind");
    require(nblocks >= 1 && nblocks <= FECX_MAX_BLOCKS && block_ne, FECX_ERR_ARG, "nblocks must be in 1..16");
    std::vector<int64_t> bs(nblocks + 1, 0);
    for (int b = 0; b < nblocks; ++b) {
      require(block_ne[b] >= 0, FECX_ERR_ARG, "negative block size");
      bs[b + 1] = bs[b] + block_ne[b];
    }
    const int64_t ne = bs[nblocks];
    require(nn > 0 && ne > 0 && coords && conn, FECX_ERR_ARG, "empty mesh");
    require(nn < (int64_t(1) << 31) && ne * ei.nnpe < (int64_t(1) << 31), FECX_ERR_ARG, "mesh exceeds int32 indexing");
    ctx->kind = elem_kind; ctx->nd = nd; ctx->nnpe = ei.nnpe; ctx->nq = ei.nq; ctx->nn = nn; ctx->ne = ne;
    ctx->conn0.resize(ne * ei.nnpe);
    for (int64_t i = 0; i < ne * ei.nnpe; ++i) {
      require(conn[i] >= 1 && conn[i] <= nn, FECX_ERR_ARG, "connectivity entry out of range");
      ctx->conn0[i] = (int32_t)(conn[i] - 1);
    }
    ctx->coords.assign(coords, coords + nn * nd);
    // coordinates as an FP32 hi/lo pair so element-local differences keep FP64-level accuracy
    std::vector<float> hi(nn * nd), lo(nn * nd);
    for (int64_t i = 0; i < nn * nd; ++i) {
      hi[i] = (float)coords[i];
      lo[i] = (float)(coords[i] - (double)hi[i]);
    }
    upload(ctx, "xhi", hi);
    upload(ctx, "xlo", lo);
    ctx->block_start = bs;
    ctx->block_mp.assign(nblocks, fecx::MatParams{});
    ctx->block_set.assign(nblocks, 0);
    ctx->has_mesh = true;
    ctx->has_dofs = ctx->has_pattern = ctx->has_material = false;
  });
}

int fecx_set_dofs(fecx_ctx* ctx, int nf, const int64_t* dirichlet, int64_t ndir) {
  return guard(ctx, [&] {
    require(ctx->has_mesh, FECX_ERR_STATE, "fecx_set_mesh has not been called");
    require(nf >= 1 && nf <= 3, FECX_ERR_ARG, "nf must be 1..3");
    require(ndir >= 0 && (ndir == 0 || dirichlet), FECX_ERR_ARG, "bad Dirichlet list");
    const int64_t N = ctx->nn * nf;
    ctx->nf = nf;
    ctx->dirichlet.assign(dirichlet, dirichlet + ndir);
    std::vector<char> isbc(N + 1, 0);
    for (int64_t i = 0; i < ndir; ++i) {
      require(dirichlet[i] >= 1 && dirichlet[i] <= N, FECX_ERR_ARG, "Dirichlet dof out of range");
      require(i == 0 || dirichlet[i] > dirichlet[i - 1], FECX_ERR_ARG, "Dirichlet dofs must be sorted and unique");
      isbc[dirichlet[i]] = 1;
    }
    ctx->unknown_dofs.clear();
    for (int64_t d = 1; d <= N; ++d)
      if (!isbc[d]) ctx->unknown_dofs.push_back(d);
    ctx->node_free.assign(ctx->nn, 1);
    ctx->partial = false;
    // condensed mode keeps every dof in the pattern; otherwise a node stays in the node-block pattern while any of its
    // dofs is free, and a node with some dofs constrained makes the dof-level compaction necessary
    for (int64_t n = 0; n < ctx->nn && !ctx->condensed; ++n) {
      int c = 0;
      for (int i = 0; i < nf; ++i) c += isbc[n * nf + i + 1];
      ctx->node_free[n] = c < nf;
      if (c > 0 && c < nf) ctx->partial = true;
    }
    std::vector<int64_t> u0(ctx->unknown_dofs.size()), d0(ndir);
    for (size_t i = 0; i < u0.size(); ++i) u0[i] = ctx->unknown_dofs[i] - 1;
    for (int64_t i = 0; i < ndir; ++i) d0[i] = dirichlet[i] - 1;
    upload(ctx, "unknown_idx", u0);
    upload(ctx, "dirichlet_idx", d0);
    ctx->has_dofs = true;
    ctx->has_pattern = false;
  });
}

int fecx_set_condensed(fecx_ctx* ctx, int on) {
  return guard(ctx, [&] {
    require(ctx->has_mesh, FECX_ERR_STATE, "fecx_set_mesh has not been called");
    ctx->condensed = on != 0;
    ctx->has_dofs = ctx->has_pattern = false;  // the DOF layout and the pattern depend on the mode
  });
}

int fecx_get_unknown_dofs(fecx_ctx* ctx, int64_t* n, int64_t* unknown) {
  return guard(ctx, [&] {
    require(ctx->has_dofs, FECX_ERR_STATE, "fecx_set_dofs has not been called");
    if (n) *n = (int64_t)ctx->unknown_dofs.size();
    if (unknown) std::memcpy(unknown, ctx->unknown_dofs.data(), ctx->unknown_dofs.size() * sizeof(int64_t));
  });
}

int fecx_build_pattern(fecx_ctx* ctx, int64_t* nnz) {
  return guard(ctx, [&] {
    require(ctx->has_dofs, FECX_ERR_STATE, "fecx_set_dofs has not been called");
    ctx->pairs = fecx::make_pairs(ctx->nnpe);
    require((ctx->ne * ctx->pairs.npairs * ctx->nf * ctx->nf) < (int64_t(1) << 31), FECX_ERR_ARG,
            "element storage exceeds 2^31 words");
    fecx::build_host_pattern(ctx->hp, ctx->nn, ctx->ne, ctx->nnpe, ctx->conn0.data(), ctx->nf, ctx->node_free,
                             ctx->pairs, ctx->nd, ctx->coords.data(), ctx->block_start);
    const int nf = ctx->nf;
    const auto& hp = ctx->hp;
    // column slots are bytes with 0xFF for "not a free node"; the reduction takes rows of up to 256 blocks
    require(hp.maxdeg <= 255, FECX_ERR_UNSUPPORTED, "a node couples to more than 255 nodes");
    ctx->nnz_full = ctx->nnz = hp.nnzb * nf * nf;
    ctx->rptr.clear();
    ctx->rcol.clear();
    ctx->red_of_full.clear();
    if (ctx->partial) {
      // the reference's reduced CSR over the unknown dofs (_update_dofs!, src/DofManagers.jl; SparsityPatterns.jl):
      // rows and columns in ascending dof order, which is node-block order restricted to the unknown dofs
      std::vector<int64_t> ridx((size_t)ctx->nn * nf, -1);
      for (size_t i = 0; i < ctx->unknown_dofs.size(); ++i) ridx[ctx->unknown_dofs[i] - 1] = (int64_t)i;
      std::vector<int64_t> cmap;
      ctx->red_of_full.assign((size_t)ctx->nnz_full, -1);
      ctx->rptr.push_back(0);
      for (int64_t r = 0; r < hp.nfree; ++r) {
        const int64_t n = hp.free_node[r], b0 = hp.bptr[r], deg = hp.bptr[r + 1] - b0;
        for (int i = 0; i < nf; ++i) {
          if (ridx[n * nf + i] < 0) continue;
          require(ridx[n * nf + i] == (int64_t)ctx->rptr.size() - 1, FECX_ERR_STATE, "unknown dof order");
          for (int64_t sl = 0; sl < deg; ++sl) {
            const int64_t m = hp.free_node[hp.bcol[b0 + sl]];
            for (int k = 0; k < nf; ++k) {
              if (ridx[m * nf + k] < 0) continue;
              const int64_t full = (int64_t)nf * nf * b0 + (int64_t)i * nf * deg + (int64_t)nf * sl + k;
              ctx->red_of_full[full] = (int64_t)cmap.size();
              cmap.push_back(full);
              ctx->rcol.push_back(ridx[m * nf + k]);
            }
          }
          ctx->rptr.push_back((int64_t)ctx->rcol.size());
        }
      }
      require((int64_t)ctx->rptr.size() == (int64_t)ctx->unknown_dofs.size() + 1, FECX_ERR_STATE, "row count");
      ctx->nnz = (int64_t)cmap.size();
      require(ctx->nnz < (int64_t(1) << 31), FECX_ERR_UNSUPPORTED, "more than 2^31 - 1 CSR entries");
      std::vector<int32_t> rcol32(ctx->rcol.begin(), ctx->rcol.end());
      upload(ctx, "cmap", cmap);
      upload(ctx, "rptr", ctx->rptr);
      upload(ctx, "rcol", rcol32);
      dev(ctx, "vals_full", sizeof(double) * (size_t)std::max<int64_t>(ctx->nnz_full, 1));
    }
    std::vector<int32_t> conn_s(ctx->ne * ctx->nnpe);  // connectivity in element storage order
    for (int64_t s = 0; s < ctx->ne; ++s)
      for (int a = 0; a < ctx->nnpe; ++a) conn_s[s * ctx->nnpe + a] = ctx->conn0[(int64_t)hp.eperm[s] * ctx->nnpe + a];
    upload(ctx, "conn", conn_s);
    upload(ctx, "eperm", hp.eperm);
    upload(ctx, "epos", hp.epos);
    {  // reduction row descriptors (16 bytes, one load instead of a chain of dependent ones) and incidence lists.
       // Scalar hex8: lists in processing order, light rows (at most 8 incidences; 8 for every interior node) in 8
       // slots each, row p at 8 p, so the reduction loads a light row's list together with its descriptor instead of
       // after it (unused slots repeat the first incidence and match no column slot); heavy rows packed after them.
       // Measured on the Poisson 128^3 reduction: -1 % (matrix), -25 % (vector). Other kinds keep the lists in node
       // order: processing order measured +0.6 % (neo-Hookean hex8 96^3) and +5 % (tet10 80^3) on the matrix reduction.
      const bool pordered = ctx->nnpe == 8 && nf == 1;
      require(hp.nnzb < (int64_t(1) << 31), FECX_ERR_UNSUPPORTED, "more than 2^31 - 1 CSR blocks");
      const int64_t nl = pordered ? hp.nlight : 0;
      const int vs = hp.vstride;
      int64_t total = nl * 8;
      if (pordered)
        for (int64_t p = nl; p < ctx->nn; ++p) total += hp.voff[hp.norder[p] + 1] - hp.voff[hp.norder[p]];
      require(total < (int64_t(1) << 31), FECX_ERR_UNSUPPORTED, "more than 2^31 - 1 incidence slots");
      std::vector<int32_t> ri((size_t)ctx->nn * 4), pv(pordered ? (size_t)total : 0);
      std::vector<uint8_t> ps(pordered ? (size_t)total * vs : 0, 0xFF);
      int64_t c = 0;
      for (int64_t p = 0; p < ctx->nn; ++p) {
        const int32_t n = hp.norder[p];
        const int32_t r = hp.free_rank[n];
        const int64_t deg = r >= 0 ? hp.bptr[r + 1] - hp.bptr[r] : 0;
        const int32_t k0 = hp.voff[n], ninc = hp.voff[n + 1] - k0;
        require(ninc < 65536, FECX_ERR_UNSUPPORTED, "a node has more than 65535 incident elements");
        ri[4 * p + 0] = n;
        ri[4 * p + 1] = pordered ? (int32_t)c : k0;
        ri[4 * p + 2] = r >= 0 ? (int32_t)hp.bptr[r] : 0;
        ri[4 * p + 3] = ninc | (int32_t)(deg << 16);
        if (!pordered) continue;
        for (int32_t t = 0; t < ninc; ++t) {
          pv[c + t] = hp.vent[k0 + t];
          std::memcpy(&ps[(size_t)(c + t) * vs], &hp.vslot[(size_t)(k0 + t) * vs], vs);
        }
        if (p < nl) {
          require(ninc <= 8, FECX_ERR_STATE, "light row with more than 8 incidences");
          for (int32_t t = ninc; t < 8; ++t) pv[c + t] = ninc > 0 ? hp.vent[k0] : 0;
          c += 8;
        } else {
          c += ninc;
        }
      }
      upload(ctx, "rowinfo", ri);
      upload(ctx, "vent", pordered ? pv : hp.vent);
      upload(ctx, "vslot", pordered ? ps : hp.vslot);
    }
    {  // tet10 vector rows: the light and the heavy chunks of one launch, in the order of their first rows' Morton
       // ranks, so that the vertex and edge rows of the same elements are reduced close together in time and share
       // the element storage lines they read in L2
      std::vector<int2> mg;
      if (ctx->nnpe == 10 && nf == 3 && hp.nlight > 0 && hp.nlight < ctx->nn) {
        const int64_t nl = hp.nlight, nh = ctx->nn - nl;
        const int64_t cl = (nl + fecx::RED_MERGE_LIGHT - 1) / fecx::RED_MERGE_LIGHT;
        const int64_t ch = (nh + fecx::RED_MERGE_HEAVY - 1) / fecx::RED_MERGE_HEAVY;
        mg.reserve(cl + ch);
        int64_t i = 0, j = 0;
        while (i < cl || j < ch) {
          const int32_t ki = i < cl ? hp.nrank[hp.norder[i * fecx::RED_MERGE_LIGHT]] : INT32_MAX;
          const int32_t kj = j < ch ? hp.nrank[hp.norder[nl + j * fecx::RED_MERGE_HEAVY]] : INT32_MAX;
          if (ki <= kj) mg.push_back(make_int2((int)(i++ * fecx::RED_MERGE_LIGHT), 0));
          else mg.push_back(make_int2((int)(nl + j++ * fecx::RED_MERGE_HEAVY), 1));
        }
      }
      upload(ctx, "rmerge", mg);
    }
    upload(ctx, "bptr", hp.bptr);
    upload(ctx, "bcol", hp.bcol);
    {  // staged row word j = (b, c) of local node a -> word inside the symmetric element storage (K_ba transposed)
      const int nnpe = ctx->nnpe, bw = nnpe * nf * nf;
      std::vector<int16_t> off((size_t)nnpe * bw);
      for (int la = 0; la < nnpe; ++la)
        for (int j = 0; j < bw; ++j) {
          const int b = j / (nf * nf), c = j % (nf * nf);
          const fecx::PairRef pr = fecx::pair_ref(nnpe, la, b);
          off[(size_t)la * bw + j] =
              (int16_t)(fecx::pair_pos(nnpe, la, b, pr.slot) * nf * nf + (pr.ab ? c : (c % nf) * nf + c / nf));
        }
      upload(ctx, "kel_off", off);
    }
    dev(ctx, "Rel", sizeof(float) * ctx->ne * ctx->nnpe * nf);
    dev(ctx, "Kel", sizeof(float) * ctx->ne * ctx->pairs.npairs * nf * nf);
    if (detect_structured_hex8(ctx) && ctx->device >= 0 && !build_nodeinfo(ctx)) ctx->structured = false;
    ctx->has_pattern = true;
    if (nnz) *nnz = ctx->nnz;
  });
}

int fecx_get_pattern(fecx_ctx* ctx, int64_t* rowptr, int64_t* colval, int base) {
  return guard(ctx, [&] {
    require(ctx->has_pattern, FECX_ERR_STATE, "fecx_build_pattern has not been called");
    require(base == 0 || base == 1, FECX_ERR_ARG, "index base must be 0 or 1");
    const auto& hp = ctx->hp;
    const int nf = ctx->nf;
    if (ctx->partial) {
      for (size_t i = 0; i < ctx->rptr.size(); ++i)
        if (rowptr) rowptr[i] = ctx->rptr[i] + base;
      for (size_t i = 0; i < ctx->rcol.size(); ++i)
        if (colval) colval[i] = ctx->rcol[i] + base;
      return;
    }
    int64_t pos = 0;
    for (int64_t r = 0; r < hp.nfree; ++r) {
      const int64_t deg = hp.bptr[r + 1] - hp.bptr[r];
      for (int i = 0; i < nf; ++i) {
        if (rowptr) rowptr[r * nf + i] = pos + base;
        for (int64_t s = 0; s < deg; ++s)
          for (int k = 0; k < nf; ++k) {
            if (colval) colval[pos] = (int64_t)nf * hp.bcol[hp.bptr[r] + s] + k + base;
            ++pos;
          }
      }
    }
    if (rowptr) rowptr[hp.nfree * nf] = pos + base;
  });
}

int fecx_get_coo_map(fecx_ctx* ctx, int64_t* map) {
  return guard(ctx, [&] {
    require(ctx->has_pattern, FECX_ERR_STATE, "fecx_build_pattern has not been called");
    require(map != nullptr, FECX_ERR_ARG, "map is NULL");
    const auto& hp = ctx->hp;
    const int nf = ctx->nf, nnpe = ctx->nnpe, NDOF = nf * nnpe;
    for (int64_t e = 0; e < ctx->ne; ++e) {
      const int32_t* c = &ctx->conn0[e * nnpe];
      for (int a = 0; a < nnpe; ++a) {
        const int32_t ra = hp.free_rank[c[a]];
        for (int i = 0; i < nf; ++i) {
          int64_t* row = map + ((e * NDOF) + a * nf + i) * NDOF;
          for (int b = 0; b < nnpe; ++b) {
            const int32_t rb = hp.free_rank[c[b]];
            int64_t base = -1;
            int64_t deg = 0;
            if (ra >= 0 && rb >= 0) {
              const int32_t* b0 = hp.bcol.data() + hp.bptr[ra];
              const int32_t* b1 = hp.bcol.data() + hp.bptr[ra + 1];
              const int64_t s = std::lower_bound(b0, b1, rb) - b0;
              deg = hp.bptr[ra + 1] - hp.bptr[ra];
              base = (int64_t)nf * nf * hp.bptr[ra] + (int64_t)i * nf * deg + (int64_t)nf * s;
            }
            for (int k = 0; k < nf; ++k) {
              int64_t v = base < 0 ? -1 : base + k;
              if (v >= 0 && ctx->partial) v = ctx->red_of_full[v];  // -1 where the row or the column is constrained
              row[b * nf + k] = v;
            }
          }
        }
      }
    }
  });
}

// Physics and properties of block `block` (every block for block = -1); (re)allocates the quadrature state, set to
// zero, when any block carries state.
static void set_material(fecx_ctx* ctx, int block, int kind, const double* props, int nprops) {
    require(ctx->has_mesh, FECX_ERR_STATE, "fecx_set_mesh has not been called");
    const int nb = (int)ctx->block_mp.size();
    require(block >= -1 && block < nb, FECX_ERR_ARG, "block index out of range");
    require(nprops >= 0 && nprops <= 8 && (nprops == 0 || props), FECX_ERR_ARG, "bad properties");
    require(kind >= FECX_POISSON && kind <= FECX_NH_SURROGATE, FECX_ERR_UNSUPPORTED, "unsupported physics kind");
    require(nb == 1 || (kind != FECX_J2_SURROGATE && kind != FECX_NH_SURROGATE), FECX_ERR_UNSUPPORTED,
            "surrogate materials need a one-block mesh");
    fecx::MatParams mp{};
    mp.phys = kind;
    for (int i = 0; i < nprops; ++i) mp.p[i] = (float)props[i];
    if (kind == FECX_POISSON) {
      require(nprops >= 2, FECX_ERR_ARG, "Poisson needs {source_kind, coefficient}");
      mp.source_kind = (int)props[0];
    }
    if (kind == FECX_LINEAR_ELASTIC) require(nprops >= 3, FECX_ERR_ARG, "linear elastic needs {rho, K, G}");
    const bool nh = kind == FECX_NEOHOOKEAN || kind == FECX_NH_SURROGATE;
    const bool j2 = kind == FECX_J2 || kind == FECX_J2_SURROGATE;
    if (nh) require(nprops >= 2, FECX_ERR_ARG, "neo-Hookean needs {lambda, mu}");
    if (j2) {
      require(nprops >= 4, FECX_ERR_ARG, "J2 needs {E, nu, sigma_y, H}");
      // the derived constants of the return mapping, in FP64 once here instead of FP32 divisions at every point
      const double E = props[0], nu = props[1], Hh = props[3];
      const double mu = E / (2 * (1 + nu)), kap = E / (3 * (1 - 2 * nu));
      mp.p[4] = (float)mu;
      mp.p[5] = (float)kap;
      mp.p[6] = (float)(1.0 / (2 * mu + 2 * Hh / 3));
      mp.p[7] = (float)(1.0 / (1 + Hh / (3 * mu)));
    }
    for (int b = 0; b < nb; ++b)
      if (block < 0 || b == block) { ctx->block_mp[b] = mp; ctx->block_set[b] = 1; }
    ctx->mp = ctx->block_mp[0];
    int ns = 0;
    bool all = true;
    for (int b = 0; b < nb; ++b) {
      const int ph = ctx->block_mp[b].phys;
      if (ctx->block_set[b] && (ph == FECX_J2 || ph == FECX_J2_SURROGATE)) ns = 7;
      all = all && ctx->block_set[b];
    }
    ctx->ns = ns;
    if (ctx->ns && ctx->device >= 0) {
      const size_t bytes = sizeof(float) * ctx->ns * ctx->nq * ctx->ne;
      HIPCHK(hipMemset(dev(ctx, "state_old", bytes), 0, bytes));
      HIPCHK(hipMemset(dev(ctx, "state_new", bytes), 0, bytes));
    }
    ctx->has_material = all;
}

int fecx_set_material(fecx_ctx* ctx, int kind, const double* props, int nprops) {
  return guard(ctx, [&] { set_material(ctx, -1, kind, props, nprops); });
}

int fecx_set_block_material(fecx_ctx* ctx, int block, int kind, const double* props, int nprops) {
  return guard(ctx, [&] {
    require(block >= 0, FECX_ERR_ARG, "block index out of range");
    set_material(ctx, block, kind, props, nprops);
  });
}

int fecx_get_num_blocks(fecx_ctx* ctx, int* nblocks) {
  return guard(ctx, [&] {
    require(ctx->has_mesh, FECX_ERR_STATE, "fecx_set_mesh has not been called");
    require(nblocks != nullptr, FECX_ERR_ARG, "nblocks is NULL");
    *nblocks = (int)ctx->block_mp.size();
  });
}

int fecx_get_state_size(fecx_ctx* ctx, int* ns, int* nq) {
  return guard(ctx, [&] {
    if (ns) *ns = ctx->ns;
    if (nq) *nq = ctx->nq;
  });
}

int fecx_set_state_old(fecx_ctx* ctx, const float* st) {
  return guard(ctx, [&] {
    check_device(ctx);
    require(ctx->ns > 0, FECX_ERR_STATE, "physics has no state");
    HIPCHK(hipMemcpy(get<float>(ctx, "state_old"), st, sizeof(float) * ctx->ns * ctx->nq * ctx->ne,
                     hipMemcpyHostToDevice));
  });
}

int fecx_get_state_new(fecx_ctx* ctx, float* st) {
  return guard(ctx, [&] {
    check_device(ctx);
    require(ctx->ns > 0, FECX_ERR_STATE, "physics has no state");
    HIPCHK(hipMemcpy(st, get<float>(ctx, "state_new"), sizeof(float) * ctx->ns * ctx->nq * ctx->ne,
                     hipMemcpyDeviceToHost));
  });
}

int fecx_state_commit(fecx_ctx* ctx) {
  return guard(ctx, [&] {
    require(ctx->ns > 0, FECX_ERR_STATE, "physics has no state");
    std::swap(ctx->buf["state_old"], ctx->buf["state_new"]);
  });
}

int fecx_assemble_vector(fecx_ctx* ctx, const double* U, double* R) {
  return guard(ctx, [&] {
    require(R != nullptr, FECX_ERR_ARG, "R is NULL");
    assemble(ctx, fecx::MODE_R, U, nullptr, R, nullptr);
  });
}

int fecx_assemble_matrix(fecx_ctx* ctx, const double* U, double* vals) {
  return guard(ctx, [&] {
    require(vals != nullptr || ctx->nnz == 0, FECX_ERR_ARG, "vals is NULL");
    assemble(ctx, fecx::MODE_K, U, nullptr, nullptr, vals);
  });
}

int fecx_assemble_vector_matrix(fecx_ctx* ctx, const double* U, double* R, double* vals) {
  return guard(ctx, [&] {
    require(R != nullptr && (vals != nullptr || ctx->nnz == 0), FECX_ERR_ARG, "R or vals is NULL");
    assemble(ctx, fecx::MODE_R | fecx::MODE_K, U, nullptr, R, vals);
  });
}

int fecx_assemble_vector_matrix_noalloc(fecx_ctx* ctx, const double* U, double* R, double* vals) {
  return guard(ctx, [&] {
    require(R != nullptr && (vals != nullptr || ctx->nnz == 0), FECX_ERR_ARG, "R or vals is NULL");
    assemble(ctx, fecx::MODE_R | fecx::MODE_K, U, nullptr, R, vals, true);
  });
}

int fecx_assemble_matrix_action(fecx_ctx* ctx, const double* U, const double* V, double* KV) {
  return guard(ctx, [&] {
    require(V != nullptr && KV != nullptr, FECX_ERR_ARG, "V or KV is NULL");
    assemble(ctx, fecx::MODE_ACTION_ASM, U, V, KV, nullptr);
  });
}

int fecx_assemble_matrix_free_action(fecx_ctx* ctx, const double* U, const double* V, double* KV) {
  return guard(ctx, [&] {
    require(V != nullptr && KV != nullptr, FECX_ERR_ARG, "V or KV is NULL");
    assemble(ctx, fecx::MODE_ACTION_MF, U, V, KV, nullptr);
  });
}

int fecx_assemble_diagonal(fecx_ctx* ctx, const double* U, double* D) {
  return guard(ctx, [&] {
    require(D != nullptr, FECX_ERR_ARG, "D is NULL");
    assemble(ctx, fecx::MODE_DIAG, U, nullptr, D, nullptr);
  });
}

int fecx_assemble_mass(fecx_ctx* ctx, double density, double* vals) {
  return guard(ctx, [&] {
    require(vals != nullptr || ctx->nnz == 0, FECX_ERR_ARG, "vals is NULL");
    assemble_mass(ctx, fecx::MASS_MATRIX, density, nullptr, nullptr, vals);
  });
}

int fecx_assemble_lumped_mass(fecx_ctx* ctx, double density, double* M) {
  return guard(ctx, [&] {
    require(M != nullptr, FECX_ERR_ARG, "M is NULL");
    assemble_mass(ctx, fecx::MASS_LUMPED, density, nullptr, M, nullptr);
  });
}

int fecx_assemble_mass_diagonal(fecx_ctx* ctx, double density, double* D) {
  return guard(ctx, [&] {
    require(D != nullptr, FECX_ERR_ARG, "D is NULL");
    assemble_mass(ctx, fecx::MASS_DIAG, density, nullptr, D, nullptr);
  });
}

int fecx_assemble_mass_action(fecx_ctx* ctx, double density, const double* V, double* MV) {
  return guard(ctx, [&] {
    require(V != nullptr && MV != nullptr, FECX_ERR_ARG, "V or MV is NULL");
    assemble_mass(ctx, fecx::MASS_ACTION, density, V, MV, nullptr);
  });
}

int fecx_update_field_unknowns(fecx_ctx* ctx, const double* Uu, double* U) {
  return guard(ctx, [&] {
    check_device(ctx);
    require(ctx->has_dofs, FECX_ERR_STATE, "fecx_set_dofs has not been called");
    const int64_t n = (int64_t)ctx->unknown_dofs.size();
    if (ctx->condensed) fecx::launch_copy_indexed(n, get<int64_t>(ctx, "unknown_idx"), Uu, U, ctx->stream);
    else fecx::launch_scatter_unknowns(n, get<int64_t>(ctx, "unknown_idx"), Uu, U, ctx->stream);
    HIPCHK(hipGetLastError());
  });
}

int fecx_update_field_dirichlet(fecx_ctx* ctx, const double* vals, double* U) {
  return guard(ctx, [&] {
    check_device(ctx);
    require(ctx->has_dofs, FECX_ERR_STATE, "fecx_set_dofs has not been called");
    fecx::launch_scatter_unknowns((int64_t)ctx->dirichlet.size(), get<int64_t>(ctx, "dirichlet_idx"), vals, U,
                                  ctx->stream);
    HIPCHK(hipGetLastError());
  });
}

int fecx_extract_field_unknowns(fecx_ctx* ctx, const double* U, double* Uu) {
  return guard(ctx, [&] {
    check_device(ctx);
    require(ctx->has_dofs, FECX_ERR_STATE, "fecx_set_dofs has not been called");
    fecx::launch_gather_unknowns((int64_t)ctx->unknown_dofs.size(), get<int64_t>(ctx, "unknown_idx"), U, Uu,
                                 ctx->stream);
    HIPCHK(hipGetLastError());
  });
}

int fecx_set_scatter(fecx_ctx* ctx, int design) {
  return guard(ctx, [&] {
    require(design >= FECX_SCATTER_FUSED && design <= FECX_SCATTER_AUTO, FECX_ERR_ARG, "unknown scatter design");
    ctx->scatter = design;
  });
}

int fecx_get_scatter_path(fecx_ctx* ctx, int* path) {
  return guard(ctx, [&] {
    require(path != nullptr, FECX_ERR_ARG, "path is NULL");
    require(ctx->has_pattern && ctx->has_material, FECX_ERR_STATE, "pattern and material must be set");
    const bool surrogate = ctx->mp.phys == FECX_J2_SURROGATE || ctx->mp.phys == FECX_NH_SURROGATE;
    *path = scatter_path(ctx, fecx::MODE_R | fecx::MODE_K, surrogate, false);
  });
}

int fecx_set_chunks(fecx_ctx* ctx, int chunks) {
  return guard(ctx, [&] {
    require(chunks >= 0 && chunks <= 1024, FECX_ERR_ARG, "chunks must be in [0, 1024]");
    ctx->chunks = chunks;
  });
}

int fecx_set_merged_rows(fecx_ctx* ctx, int on) {
  return guard(ctx, [&] { ctx->merged_rows = on != 0; });
}

int fecx_set_fp64_geometry(fecx_ctx* ctx, int on) {
  return guard(ctx, [&] { ctx->fp64_geometry = on != 0; });
}

int fecx_set_profiling(fecx_ctx* ctx, int on) {
  return guard(ctx, [&] { ctx->profiling = on != 0; });
}

int fecx_profile_count(fecx_ctx* ctx, int* count) {
  return guard(ctx, [&] { *count = (int)ctx->prof_names.size(); });
}

int fecx_profile_entry(fecx_ctx* ctx, int i, char* name, int name_len, int64_t* launches, double* total_ms) {
  return guard(ctx, [&] {
    require(i >= 0 && i < (int)ctx->prof_names.size(), FECX_ERR_ARG, "profile index out of range");
    const std::string& n = ctx->prof_names[i];
    if (name && name_len > 0) {
      std::strncpy(name, n.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    const ProfEntry& pe = ctx->prof[n];
    if (launches) *launches = pe.launches;
    if (total_ms) *total_ms = pe.ms;
  });
}

int fecx_profile_reset(fecx_ctx* ctx) {
  return guard(ctx, [&] {
    ctx->prof.clear();
    ctx->prof_names.clear();
  });
}

}  // extern "C"
