// Host-side sparsity pattern and COO->CSR index maps (built once per DOF layout).
//
// The reference builds Is/Js/unknown_dofs over every element (i outer, j inner), stable-sorts the
// packed key (I<<32)|J into `permutation`, and on every stiffness() call reduces the COO storage
// through SparseArrays.sparse! (src/assemblers/SparsityPatterns.jl:53-117, 160-231, 301-329).
// Here the same reduction is precomputed in block form: the sorted unique (I, J) set is the CSR
// pattern, and for every CSR block the list of element-storage blocks that sum into it (the runs of
// equal keys in `permutation`, in ascending element order) is the inverse COO->CSR map.
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

void build_host_pattern(HostPattern& hp, int64_t nn, int64_t ne, int nnpe, const int32_t* conn0, int nf,
                        const std::vector<char>& node_free, const PairTable& pt) {
  hp = HostPattern();
  hp.nf = nf; hp.nnpe = nnpe; hp.nn = nn; hp.ne = ne;
  // node -> (element, local node), ascending element order (counting sort over elements in order)
  hp.voff.assign(nn + 1, 0);
  for (int64_t i = 0; i < ne * nnpe; ++i) hp.voff[conn0[i] + 1]++;
  for (int64_t n = 0; n < nn; ++n) hp.voff[n + 1] += hp.voff[n];
  hp.vent.resize(ne * nnpe);
  {
    std::vector<int32_t> fill(hp.voff.begin(), hp.voff.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < nnpe; ++a) hp.vent[fill[conn0[e * nnpe + a]]++] = (int32_t)(e * nnpe + a);
  }
  // free ranks
  hp.free_rank.assign(nn, -1);
  for (int64_t n = 0; n < nn; ++n)
    if (node_free[n]) { hp.free_rank[n] = (int32_t)hp.nfree; hp.free_node.push_back((int32_t)n); ++hp.nfree; }
  const int64_t nfree = hp.nfree;
  // pass 1: degree (distinct free neighbour nodes) and contribution count per free row
  std::vector<int64_t> deg(nfree), ctot(nfree);
  parallel_for(nfree, [&](int64_t b, int64_t e, int) {
    std::vector<int32_t> nb;
    for (int64_t r = b; r < e; ++r) {
      int32_t n = hp.free_node[r];
      nb.clear();
      int64_t c = 0;
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        int64_t el = hp.vent[k] / nnpe;
        for (int bb = 0; bb < nnpe; ++bb) {
          int32_t m = conn0[el * nnpe + bb];
          if (hp.free_rank[m] >= 0) { nb.push_back(hp.free_rank[m]); ++c; }
        }
      }
      std::sort(nb.begin(), nb.end());
      deg[r] = std::unique(nb.begin(), nb.end()) - nb.begin();
      ctot[r] = c;
    }
  });
  hp.bptr.assign(nfree + 1, 0);
  std::vector<int64_t> coff(nfree + 1, 0);
  for (int64_t r = 0; r < nfree; ++r) { hp.bptr[r + 1] = hp.bptr[r] + deg[r]; coff[r + 1] = coff[r] + ctot[r]; }
  hp.nnzb = hp.bptr[nfree];
  hp.ncontrib = coff[nfree];
  hp.bcol.resize(hp.nnzb);
  hp.boff.assign(hp.nnzb + 1, 0);
  hp.bent.resize(hp.ncontrib);
  // pass 2: columns, per-block contribution lists in ascending element order
  parallel_for(nfree, [&](int64_t b, int64_t e, int) {
    std::vector<int32_t> nb;
    std::vector<int64_t> cnt;
    for (int64_t r = b; r < e; ++r) {
      int32_t n = hp.free_node[r];
      nb.clear();
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        int64_t el = hp.vent[k] / nnpe;
        for (int bb = 0; bb < nnpe; ++bb) {
          int32_t m = conn0[el * nnpe + bb];
          if (hp.free_rank[m] >= 0) nb.push_back(hp.free_rank[m]);
        }
      }
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
      int64_t b0 = hp.bptr[r];
      std::copy(nb.begin(), nb.end(), hp.bcol.begin() + b0);
      cnt.assign(nb.size() + 1, 0);
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        int64_t el = hp.vent[k] / nnpe;
        for (int bb = 0; bb < nnpe; ++bb) {
          int32_t m = conn0[el * nnpe + bb];
          if (hp.free_rank[m] < 0) continue;
          int64_t s = std::lower_bound(nb.begin(), nb.end(), hp.free_rank[m]) - nb.begin();
          cnt[s + 1]++;
        }
      }
      for (size_t s = 0; s < nb.size(); ++s) cnt[s + 1] += cnt[s];
      for (size_t s = 0; s < nb.size(); ++s) hp.boff[b0 + s + 1] = coff[r] + cnt[s + 1];
      if (r == 0) hp.boff[0] = 0;
      std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
      for (int32_t k = hp.voff[n]; k < hp.voff[n + 1]; ++k) {
        int64_t el = hp.vent[k] / nnpe;
        int a = hp.vent[k] % nnpe;
        for (int bb = 0; bb < nnpe; ++bb) {
          int32_t m = conn0[el * nnpe + bb];
          if (hp.free_rank[m] < 0) continue;
          int64_t s = std::lower_bound(nb.begin(), nb.end(), hp.free_rank[m]) - nb.begin();
          uint32_t pid = (uint32_t)(el * pt.npairs + pt.slot_of[a][bb]);
          uint32_t tr = pt.stored_ab[a][bb] ? 0u : 1u;
          hp.bent[coff[r] + pos[s]++] = pid | (tr << 31);
        }
      }
    }
  });
}

}  // namespace fecx
