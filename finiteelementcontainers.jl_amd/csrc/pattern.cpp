// Host-side sparsity pattern and COO->CSR index maps (built once per DOF layout).
//
// The reference builds Is/Js/unknown_dofs over every element (i outer, j inner), stable-sorts the
// packed key (I<<32)|J into `permutation`, and on every stiffness() call reduces the COO storage
// through SparseArrays.sparse! (src/assemblers/SparsityPatterns.jl:53-117, 160-231, 301-329).
// Here the same reduction is precomputed in block form: the sorted unique (I, J) set is the CSR
// pattern, and the inverse COO->CSR map is stored per CSR row-node: its incident (element, local node)
// pairs in ascending element order (the runs of equal keys in `permutation`, summed in that order) and,
// per incidence, the column slot each local node lands in.
#include <algorithm>
#include <numeric>
#include <thread>

#include "fecx_internal.h"

namespace fecx {

ElemInfo elem_info(int kind) {
  switch (kind) {
    case FECX_QUAD4: return {FECX_QUAD4, 4, 2, 4, 4, 10};
    case FECX_HEX8: return {FECX_HEX8, 8, 3, 8, 8, 36};
    case FECX_TET10: return {FECX_TET10, 10, 3, 4, 16, 55};
    default: return {-1, 0, 0, 0, 0, 0};
  }
}

PairTable make_pairs(int nnpe) {
  PairTable t;
  t.nnpe = nnpe;
  for (int a = 0; a < 16; ++a)
    for (int b = 0; b < 16; ++b) { t.slot_of[a][b] = -1; t.stored_ab[a][b] = false; }
  int slot = 0;
  for (int a = 0; a < nnpe; ++a) {
    t.lane_off[a] = slot;
    int dmax = (nnpe - 1) / 2;
    for (int d = 0; d <= dmax + 1; ++d) {
      if (d == dmax + 1 && !(nnpe % 2 == 0 && a < nnpe / 2)) break;
      int b = (a + d) % nnpe;
      t.pa[slot] = a;
      t.pb[slot] = b;
      t.slot_of[a][b] = slot; t.stored_ab[a][b] = true;
      t.slot_of[b][a] = slot; t.stored_ab[b][a] = (a == b);
      ++slot;
    }
    t.lane_cnt[a] = slot - t.lane_off[a];
  }
  t.lane_off[nnpe] = slot;
  t.npairs = slot;
  return t;
}

template <class F>
static void parallel_for(int64_t n, F f) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
  if (n < 4096 || nt == 1) {
    f((int64_t)0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { f(b, e, t); });
  }
  for (auto& x : th) x.join();
}

static uint64_t spread3(uint64_t x) {  // 21 bits -> every third bit
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

void build_host_pattern(HostPattern& hp, int64_t nn, int64_t ne, int nnpe, const int32_t* conn0, int nf,
                        const std::vector<char>& node_free, const PairTable& pt, int nd, const double* coords,
                        const std::vector<int64_t>& block_start) {
  (void)pt;
  hp = HostPattern();
  hp.nf = nf; hp.nnpe = nnpe; hp.nn = nn; hp.ne = ne;
  hp.vstride = (nnpe + 3) / 4 * 4;
  // Morton rank of every node (21 bits per coordinate over the bounding box), ties by node id
  std::vector<int32_t>& nrank = hp.nrank;
  nrank.assign(nn, 0);
  {
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    for (int d = 0; d < nd; ++d) {
      lo[d] = hi[d] = coords[d];
      for (int64_t n = 0; n < nn; ++n) { lo[d] = std::min(lo[d], coords[n * nd + d]); hi[d] = std::max(hi[d], coords[n * nd + d]); }
    }
    std::vector<std::pair<uint64_t, int32_t>> keyed(nn);
    parallel_for(nn, [&](int64_t b, int64_t e, int) {
      for (int64_t n = b; n < e; ++n) {
        uint64_t code = 0;
        for (int d = 0; d < nd; ++d) {
          const double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
          const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, (coords[n * nd + d] - lo[d]) / span * 2097151.0));
          code |= spread3(q) << d;
        }
        keyed[n] = {code, (int32_t)n};
      }
    });
    std::sort(keyed.begin(), keyed.end());
    hp.norder.resize(nn);
    for (int64_t i = 0; i < nn; ++i) { hp.norder[i] = keyed[i].second; nrank[keyed[i].second] = (int32_t)i; }
  }
  // element storage order: by smallest Morton rank of the element's nodes, ties by element id
  std::vector<int32_t> key(ne);
  parallel_for(ne, [&](int64_t b, int64_t e, int) {
    for (int64_t el = b; el < e; ++el) {
      int32_t m = nrank[conn0[el * nnpe]];
      for (int a = 1; a < nnpe; ++a) m = std::min(m, nrank[conn0[el * nnpe + a]]);
      key[el] = m;
    }
  });
  // Several element blocks (elements numbered block after block, as the blocks' connectivities are concatenated):
  // each block takes a contiguous range of storage positions, in block order, sorted by that key within the block,
  // so that one element-kernel launch per block runs that block's material over its range.
  {
    const int nb = (int)block_start.size() - 1;
    hp.eperm.resize(ne);
    hp.epos.resize(ne);
    hp.block_pos.assign(block_start.begin(), block_start.end());
    std::vector<int64_t> cnt(nn + 1);
    for (int b = 0; b < nb; ++b) {
      const int64_t e0 = block_start[b], e1 = block_start[b + 1];
      std::fill(cnt.begin(), cnt.end(), 0);
      cnt[0] = e0;
      for (int64_t el = e0; el < e1; ++el) cnt[key[el] + 1]++;
      for (int64_t n = 0; n < nn; ++n) cnt[n + 1] += cnt[n];
      if (nb == 1) hp.ekey = cnt;  // piece boundaries of pipelined assembly (one block only)
      for (int64_t el = e0; el < e1; ++el) {
        const int64_t s = cnt[key[el]]++;
        hp.eperm[s] = (int32_t)el;
        hp.epos[el] = (int32_t)s;
      }
    }
  }
  // node -> incidences, ascending element id (counting sort over elements in order)
  hp.voff.assign(nn + 1, 0);
  for (int64_t i = 0; i < ne * nnpe; ++i) hp.voff[conn0[i] + 1]++;
  for (int64_t n = 0; n < nn; ++n) hp.voff[n + 1] += hp.voff[n];
  std::vector<int32_t> inc_el(ne * nnpe);
  hp.vent.resize(ne * nnpe);
  {
    std::vector<int32_t> fill(hp.voff.begin(), hp.voff.end() - 1);
    for (int64_t el = 0; el < ne; ++el)
      for (int a = 0; a < nnpe; ++a) {
        const int32_t k = fill[conn0[el * nnpe + a]]++;
        inc_el[k] = (int32_t)el;
        hp.vent[k] = (int32_t)((int64_t)hp.epos[el] * nnpe + a);
      }
  }
  // free ranks
  hp.free_rank.assign(nn, -1);
  for (int64_t n = 0; n < nn; ++n)
    if (node_free[n]) { hp.free_rank[n] = (int32_t)hp.nfree; hp.free_node.push_back((int32_t)n); ++hp.nfree; }
  const int64_t nfree = hp.nfree;
  // pass 1: degree of every free row (distinct free neighbour nodes)
  std::vector<int64_t> deg(nfree);
  parallel_for(nfree, [&](int64_t b, int64_t e, int) {
    std::vector<int32_t> nb;
    for (int64_t r = b; r < e; ++r) {
      const int32_t n = hp.free_node[r];
      nb.clear();
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        const int64_t el = inc_el[k];
        for (int bb = 0; bb < nnpe; ++bb) {
          const int32_t m = conn0[el * nnpe + bb];
          if (hp.free_rank[m] >= 0) nb.push_back(hp.free_rank[m]);
        }
      }
      std::sort(nb.begin(), nb.end());
      deg[r] = std::unique(nb.begin(), nb.end()) - nb.begin();
    }
  });
  hp.bptr.assign(nfree + 1, 0);
  for (int64_t r = 0; r < nfree; ++r) {
    hp.bptr[r + 1] = hp.bptr[r] + deg[r];
    hp.maxdeg = std::max<int>(hp.maxdeg, (int)deg[r]);
  }
  hp.nnzb = hp.bptr[nfree];
  // processing order: light rows (at most 8 incidences, one reduction pass, and at most 32 column blocks) first, then
  // the heavy ones, each bucket in Morton order
  {
    std::vector<int32_t> light, heavy;
    light.reserve(nn);
    int mdl = 0;
    for (int64_t i = 0; i < nn; ++i) {
      const int32_t n = hp.norder[i];
      const int ninc = hp.voff[n + 1] - hp.voff[n];
      const int32_t r = hp.free_rank[n];
      const int d = r >= 0 ? (int)deg[r] : 0;
      if (ninc <= 8 && d <= 32) {
        light.push_back(n);
        mdl = std::max(mdl, d);
      } else {
        heavy.push_back(n);
      }
    }
    hp.nlight = (int64_t)light.size();
    hp.maxdeg_light = mdl;
    std::copy(light.begin(), light.end(), hp.norder.begin());
    std::copy(heavy.begin(), heavy.end(), hp.norder.begin() + light.size());
  }
  hp.bcol.resize(hp.nnzb);
  hp.vslot.assign((size_t)ne * nnpe * hp.vstride, 0xFF);
  // pass 2: columns and the per-incidence column slots
  parallel_for(nfree, [&](int64_t b, int64_t e, int) {
    std::vector<int32_t> nb;
    for (int64_t r = b; r < e; ++r) {
      const int32_t n = hp.free_node[r];
      nb.clear();
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        const int64_t el = inc_el[k];
        for (int bb = 0; bb < nnpe; ++bb) {
          const int32_t m = conn0[el * nnpe + bb];
          if (hp.free_rank[m] >= 0) nb.push_back(hp.free_rank[m]);
        }
      }
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
      std::copy(nb.begin(), nb.end(), hp.bcol.begin() + hp.bptr[r]);
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        const int64_t el = inc_el[k];
        uint8_t* row = &hp.vslot[(size_t)k * hp.vstride];
        for (int bb = 0; bb < nnpe; ++bb) {
          const int32_t fr = hp.free_rank[conn0[el * nnpe + bb]];
          if (fr < 0) continue;
          row[bb] = (uint8_t)(std::lower_bound(nb.begin(), nb.end(), fr) - nb.begin());
        }
      }
    }
  });
}

}  // namespace fecx
