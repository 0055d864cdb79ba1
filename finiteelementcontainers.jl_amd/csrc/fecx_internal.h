// Internal declarations shared by the host C-ABI layer and the HIP kernels of libfecx.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/fecx.h"

namespace fecx {

// ------------------------------------------------------------------ element geometry
// Lanes per element in the element kernels: hex8 and quad4 map one lane to each local node
// and each quadrature point (8 and 4); tet10 uses 16 lanes (10 nodes, 4 quadrature points).
struct ElemInfo {
  int kind, nnpe, nd, nq, lanes, npairs;
};
ElemInfo elem_info(int kind);

// Symmetric element-matrix storage: one NF x NF block per unordered node pair {a, b}. Lane a owns
// the pairs (a, (a+d) mod nnpe) for d = 0..(nnpe-1)/2, plus d = nnpe/2 when nnpe is even and
// a < nnpe/2, so every lane's blocks are contiguous in the element's storage (slot = lane-major).
struct PairTable {
  int nnpe = 0, npairs = 0;
  int lane_off[17] = {0}, lane_cnt[16] = {0};
  int pa[64] = {0}, pb[64] = {0};          // slot -> (a, b): the stored block is K_ab
  int slot_of[16][16];                     // unordered pair -> slot
  bool stored_ab[16][16];                  // true if slot_of[a][b] stores K_ab (else K_ba)
};
PairTable make_pairs(int nnpe);

// Pair (a, b) of an element -> its block in the symmetric element storage and whether that block is K_ab (true)
// or K_ba (false). Same rule as make_pairs, usable in kernels.
struct PairRef {
  int slot;
  bool ab;
};
__host__ __device__ constexpr int pair_lane_off(int nnpe, int a) {
  return a * ((nnpe - 1) / 2 + 1) + ((nnpe % 2 == 0) ? (a < nnpe / 2 ? a : nnpe / 2) : 0);
}
__host__ __device__ constexpr PairRef pair_ref(int nnpe, int a, int b) {
  const int dmax = (nnpe - 1) / 2;
  const int d1 = (b - a + nnpe) % nnpe;
  if (d1 <= dmax) return {pair_lane_off(nnpe, a) + d1, true};
  const int d2 = (a - b + nnpe) % nnpe;
  if (d2 <= dmax) return {pair_lane_off(nnpe, b) + d2, false};
  const int lo = a < b ? a : b;  // d1 == d2 == nnpe / 2
  return {pair_lane_off(nnpe, lo) + nnpe / 2, lo == a};
}

// Position of the block of the pair {a, b} (slot `slot` of the rule above) in an element's storage. tet10 (local
// nodes 0-3 vertices, 4-9 edge midpoints): the edge-edge pairs, then the vertex-edge ones, then the vertex-vertex ones,
// so that the rows of each node kind -- the reduction's two row buckets, launched one after the other -- read one
// contiguous region of the record (blocks 0-44 and 21-54) instead of blocks spread over all of it. Other kinds: the
// slot itself.
__host__ __device__ constexpr int pair_pos(int nnpe, int a, int b, int slot) {
  if (nnpe != 10) return slot;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  if (lo >= 4) {
    const int i = lo - 4, j = hi - 4;
    return i * 6 - i * (i - 1) / 2 + (j - i);
  }
  if (hi >= 4) return 21 + lo * 6 + (hi - 4);
  return 45 + lo * 4 - lo * (lo - 1) / 2 + (hi - lo);
}

// ------------------------------------------------------------------ host pattern
// Block-CSR view of the reduced CSR pattern: a "free node" has all NF dofs unknown. Free nodes are
// ranked in node order; the unknown index of (free rank r, dof i) is NF*r + i, so every CSR row of
// node r has NF*deg(r) columns made of deg(r) runs of NF consecutive columns.
//
// Processing order: nodes are ranked along a Morton (Z-order) curve of their coordinates; the reduction walks nodes
// in that order and elements are stored sorted by the smallest Morton rank among their nodes (ties by element
// id), so that a group of nodes reduced together reads a compact range of the element-private storage.
// Inverse COO->CSR map, per node n: its incident (element, local node a) pairs in ascending element id
// (the order SparseArrays.sparse! sums duplicates in), and for each incidence one byte per local node b
// giving the column slot of node conn[e][b] in row n (0xFF: not a free node).
struct HostPattern {
  int nf = 0, nnpe = 0, vstride = 0, maxdeg = 0;
  int64_t nn = 0, ne = 0, nfree = 0, nnzb = 0;
  std::vector<int32_t> free_rank;   // nn: rank or -1
  std::vector<int32_t> free_node;   // nfree: node id (0-based)
  std::vector<int64_t> bptr;        // nfree + 1 : block row pointer
  std::vector<int32_t> bcol;        // nnzb: free rank of the column node, ascending per row
  std::vector<int32_t> norder;      // processing position -> node: light rows, then heavy rows, Morton order in each
  std::vector<int32_t> nrank;       // node -> rank along the Morton curve
  int64_t nlight = 0;               // light rows: at most 8 incidences and 32 column blocks
  int maxdeg_light = 0;
  std::vector<int32_t> eperm;       // storage position -> element
  std::vector<int32_t> epos;        // element -> storage position
  std::vector<int64_t> ekey;        // nn + 1: storage positions of the elements whose smallest node rank is < p
                                    // (one element block only; empty otherwise)
  std::vector<int64_t> block_pos;   // nblocks + 1: storage position range of each element block
  std::vector<int32_t> voff;        // nn + 1 : node -> incidences
  std::vector<int32_t> vent;        // storage position * nnpe + a, ascending element id per node
  std::vector<uint8_t> vslot;       // incidences x vstride bytes: column slot per local node b
};

// conn0: 0-based node ids, ne x nnpe row-major (element-major); node_free: nn flags; coords: nd x nn (FP64);
// block_start: nblocks + 1 element-id offsets of the element blocks (blocks numbered one after the other).
void build_host_pattern(HostPattern& hp, int64_t nn, int64_t ne, int nnpe, const int32_t* conn0, int nf,
                        const std::vector<char>& node_free, const PairTable& pt, int nd, const double* coords,
                        const std::vector<int64_t>& block_start);

// ------------------------------------------------------------------ kernels
// Work-groups are dealt to the 8 XCDs round-robin; remap block index wg so that each XCD walks a contiguous range of
// work (rows or elements), which keeps the data shared by neighbouring rows or elements in that XCD's L2.
__device__ __forceinline__ int64_t xcd_chunk(int64_t wg, int64_t nwg) {
  const int64_t per = nwg / 8;  // nwg is a multiple of 8
  return (wg % 8) * per + wg / 8;
}

struct MatParams {
  int phys;
  int source_kind;
  float p[8];
};

enum Mode { MODE_R = 1, MODE_K = 2, MODE_ACTION_MF = 4, MODE_ACTION_ASM = 8, MODE_QPGRAD = 16, MODE_DIAG = 32 };

// mass-type assemblies (mass.hip): consistent mass blocks into Kel; lumped mass, mass diagonal or mass action into Rel
enum MassWhat { MASS_MATRIX = 1, MASS_LUMPED = 2, MASS_DIAG = 3, MASS_ACTION = 4 };

struct ElementArgs {
  int64_t ebeg, ne;         // element storage positions [ebeg, ne)
  const int32_t* conn;      // ne x nnpe (0-based), in element storage order
  const int32_t* eperm;     // storage position -> element id (quadrature state index)
  const float* xhi;         // nd x nn
  const float* xlo;         // nd x nn
  const double* U;          // nf x nn
  const double* V;          // nf x nn (actions)
  const float* st_old;      // ns x nq x ne
  float* st_new;
  float* Rel;               // ne x ndof   (residual or action output)
  float* Kel;               // ne x npairs x nf*nf
  int* flag;                // error flag (bit 1: bad Jacobian / NaN)
  float* hq;                // MODE_QPGRAD output: ne x nq x (nf*nd) gradients, element storage order
  const float* qpmat;       // surrogate materials: ne x nq x (nf*nd + NC) stress and tangent data
  MatParams mp;
  int fp64_geometry;        // tet10: element geometry, field gradient and stress in FP64 (fecx_set_fp64_geometry)
};

// returns 0, or -1 if the (kind, nf, phys) combination has no kernel
int launch_element(int kind, int nf, int mode, const ElementArgs& a, hipStream_t s);

// Fused residual + CSR assembly of a structured hex8 mesh (element.hip, fused_hex8_kernel).
struct FusedArgs {
  int nx, ny, nz;                // node counts per direction
  int tiles_x, tiles_y;          // node tiles per direction (set by launch_fused_hex8)
  const int32_t* epos;           // element id -> storage position
  const int2* nodeinfo;          // per node: {first CSR block of its row, mask of the 27 neighbours in the row}
  double* vals;                  // CSR values (or nullptr)
  double* R;                     // full NF x NN residual field (or nullptr)
};
// returns 0, or -1 if the (nf, phys) combination has no fused kernel (or the mesh exceeds 2^31 tile planes)
int launch_fused_hex8(int nf, const ElementArgs& a, const FusedArgs& f, hipStream_t s);
// the atomic (scatter 1) and coloured (scatter 2) designs on structured hex8 meshes; vals and R zero-filled first
int launch_scatter_hex8(int nf, int scatter, const ElementArgs& a, const FusedArgs& f, hipStream_t s);
// returns 0, or -1 if the (kind, nf) combination has no mass kernel; rho multiplies every weight
int launch_mass(int kind, int nf, int what, const ElementArgs& a, float rho, hipStream_t s);

struct ReduceArgs {
  int64_t p0, p1;           // processing positions [p0, p1) of this launch
  int64_t nlight;           // positions [0, nlight) hold the light rows (at most 8 incidences and 32 blocks)
  int maxdeg_light;         // largest row degree among them
  int onepass;              // every row of this launch has at most CH incidences (one pass, no pass-count vote)
  int64_t npad;             // positions [0, npad) hold padded rows: incidence slots vent[8 p .. 8 p + 7]
  int padded;               // this launch reads the padded slots directly (every row of it is padded)
  const int4* rowinfo;      // per processing position: {node, first incidence slot, first CSR block or 0, ninc | degree << 16}
  int nnpe, maxdeg;
  const int32_t* vent;
  const uint8_t* vslot;
  const int16_t* kel_off;   // nnpe x (nnpe*nf*nf): word of (local node a, staged word j) in an element's storage
  // tet10 rows of 3 fields, both buckets in one launch (reduce_merged_kernel): per work-group {first position, bucket},
  // light chunks of RED_MERGE_LIGHT rows and heavy chunks of RED_MERGE_HEAVY rows, merged along the Morton curve
  const int2* merged;
  int64_t nmerged;
  const float* Kel;
  const float* Rel;
  double* vals;             // CSR values, or nullptr
  double* R;                // full NF x NN field, or nullptr
};
int launch_reduce(int nf, const ReduceArgs& a, hipStream_t s);
constexpr int RED_MERGE_LIGHT = 8, RED_MERGE_HEAVY = 4;  // rows per work-group of the merged launch's two buckets

// Allocation-free assembly (noalloc.hip): a file-scope device array of NOALLOC_SCRATCH_BYTES per GPU holds the
// element-private outputs. noalloc_setup looks up the array's device address on `device` (which must be current) once;
// noalloc_scratch returns it (nullptr before setup).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(FECX_NOALLOC_SCRATCH_BYTES);
constexpr int NOALLOC_MAX_DEVICES = 64;
int noalloc_setup(int device);
void* noalloc_scratch(int device);

// DOF plumbing
void launch_scatter_unknowns(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s);
void launch_gather_unknowns(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s);
// dst[idx[i]] = src[idx[i]] (condensed mode: full-length unknown vectors)
void launch_copy_indexed(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s);

}  // namespace fecx

// ------------------------------------------------------------------ context
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct ProfEntry {
  int64_t launches = 0;
  double ms = 0;
};

struct fecx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // mesh
  bool has_mesh = false;
  int kind = -1, nd = 0, nnpe = 0, nq = 0;
  int64_t nn = 0, ne = 0;
  std::vector<int64_t> block_start;  // nblocks + 1 element-id offsets (one block: {0, ne})
  std::vector<int32_t> conn0;  // ne x nnpe, 0-based
  std::vector<double> coords;  // nd x nn
  // dofs
  bool has_dofs = false;
  bool condensed = false;             // use_condensed: every dof stays in the pattern (fecx_set_condensed)
  int nf = 0;
  std::vector<int64_t> dirichlet;     // 1-based
  std::vector<int64_t> unknown_dofs;  // 1-based
  std::vector<char> node_free;
  // structured hex8 mesh (StructuredMesh numbering): node counts per direction, and whether matrix assemblies take the
  // scatter design (fecx_set_scatter)
  int snx = 0, sny = 0, snz = 0;
  bool structured = false;
  int scatter = FECX_SCATTER_AUTO;  // fecx_set_scatter
  // pattern
  bool has_pattern = false;
  fecx::PairTable pairs;
  fecx::HostPattern hp;
  int64_t nnz = 0;       // entries of the CSR the caller sees
  int64_t nnz_full = 0;  // entries of the node-block layout the kernels assemble (NF^2 per block)
  // a Dirichlet set that constrains some but not all dofs of a node (nf > 1, not condensed): the kernels assemble the
  // node-block pattern of the nodes with a free dof into "vals_full", and a gather through "cmap" forms the reference's
  // dof-level CSR (rows and columns of the unknown dofs); rptr/rcol are that CSR (0-based), red_of_full its inverse map
  bool partial = false;
  std::vector<int64_t> rptr, rcol, red_of_full;
  // material: one per element block (foreach_block's per-block physics and properties); mp is block 0's
  bool has_material = false;
  fecx::MatParams mp{};
  std::vector<fecx::MatParams> block_mp;
  std::vector<char> block_set;
  int ns = 0;
  // device buffers
  std::map<std::string, DevBuf> buf;
  // learned surrogate (FECX_MLP_*): parameters and packed weights live in buf ("sg_*")
  int sg_arch = -1;
  float sg_tol = 0.02f;
  int64_t sg_step = 0;                          // Adam steps taken
  int64_t sg_fallback = 0, sg_points = 0;       // last surrogate pass
  bool sg_enabled = true;                       // off: surrogate physics run the plain model in the element kernel
  bool sg_freeze = false;                       // on: reuse the last pass's accept/fallback decisions
  bool sg_frozen_valid = false;
  // pipelined assembly: element chunks on stream2, reduction chunks on the context's stream
  int chunks = 0;  // 0: chosen from the element storage size
  bool merged_rows = true;  // tet10 3-field reduction: both row buckets in one launch (fecx_set_merged_rows)
  bool fp64_geometry = false;  // tet10 element geometry, field gradient and stress in FP64 (fecx_set_fp64_geometry)
  hipStream_t stream2 = nullptr;
  std::vector<hipEvent_t> events;
  // numeric error flag: device word set by the kernels, read back through a pinned host word after every assembly;
  // cleared on the device only when the last read-back did not find it clear (an error, or a call that stopped early)
  int* flag_host = nullptr;
  bool flag_dirty = true;
  // profiling
  bool profiling = false;
  std::vector<std::string> prof_names;
  std::map<std::string, ProfEntry> prof;
};
