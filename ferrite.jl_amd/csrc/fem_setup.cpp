// Native host-side setup of the assembly path: dof distribution, sparsity
// pattern and cell coloring, with the same results as Ferrite's own setup so
// that the device tables built from them match the reference bit for bit.
// (A Julia caller runs Ferrite's setup itself and hands the arrays to the
// upload entry points; these routines serve non-Julia hosts.)
//
//   fem_setup_distribute_dofs  DofHandler.jl:493-794 (first-visit numbering:
//                              cells ascending, per field vertices -> edges
//                              (key = sorted node pair) -> faces (key = three
//                              smallest nodes) -> volume; n_copies interleave)
//   fem_setup_pattern          sparsity_pattern.jl:223-253, :951-1016 (row r =
//                              diagonal + dofs of every cell containing r;
//                              rowval ascending per column)
//   fem_setup_coloring         coloring.jl:47-317 (Greedy and WorkStream)
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ferrite_mi355x_setup.h"

namespace {

struct Shape {
  int rdim, nv;
  std::vector<std::pair<int, int>> edges;   // 0-based local vertex pairs
  std::vector<std::vector<int>> faces;      // 0-based local vertex lists
};

Shape make_shape(int refshape) {
  Shape s{};
  if (refshape == FEM_REF_LINE) {
    s.rdim = 1; s.nv = 2;
    s.edges = {{0, 1}};
  } else if (refshape == FEM_REF_QUADRILATERAL) {
    s.rdim = 2; s.nv = 4;
    s.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    s.faces = {{0, 1, 2, 3}};
  } else {
    s.rdim = 3; s.nv = 8;
    s.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    s.faces = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
  }
  return s;
}

// Number of interior dofs per entity for Lagrange order 1/2 (interpolations.jl:584-667, 1075-1085).
struct EntityDofs {
  int vertex, edge, face, volume;
};

EntityDofs entity_dofs(int refshape, int order) {
  if (order == 1) return {1, 0, 0, 0};
  if (refshape == FEM_REF_LINE) return {1, 1, 0, 0};
  if (refshape == FEM_REF_QUADRILATERAL) return {1, 1, 1, 0};
  return {1, 1, 1, 1};
}

struct PairHash {
  size_t operator()(const std::pair<int64_t, int64_t>& p) const {
    return std::hash<int64_t>()(p.first * 0x9E3779B97F4A7C15ULL ^ (p.second + 0x632BE59BD9B4E019ULL));
  }
};
struct TripleKey {
  int64_t a, b, c;
  bool operator==(const TripleKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct TripleHash {
  size_t operator()(const TripleKey& k) const {
    uint64_t h = (uint64_t)k.a * 0x9E3779B97F4A7C15ULL;
    h ^= (uint64_t)k.b + 0x632BE59BD9B4E019ULL + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.c + 0x94D049BB133111EBULL + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

int hw_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return (int)std::min(n, 16u);
}

template <class F>
void parallel_for(int64_t n, F&& f) {
  const int nt = hw_threads();
  if (n < 4096 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([&f, lo, hi] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int32_t fem_setup_distribute_dofs(int32_t refshape, int64_t ncells, int32_t npc, const int64_t* cell_nodes,
                                  int64_t nnodes, int32_t nfields, const int32_t* field_order,
                                  const int32_t* field_vdim, int64_t* cell_dofs, int32_t* ndofs_per_cell,
                                  int64_t* ndofs) {
  if (!cell_nodes || !field_order || !field_vdim || !ndofs || !ndofs_per_cell || ncells <= 0 || nfields <= 0)
    return FEM_ERR_ARG;
  if (refshape != FEM_REF_LINE && refshape != FEM_REF_QUADRILATERAL && refshape != FEM_REF_HEXAHEDRON)
    return FEM_ERR_ARG;
  const Shape sh = make_shape(refshape);
  if (npc < sh.nv) return FEM_ERR_ARG;
  // per-cell dof count
  int ndpc = 0;
  std::vector<EntityDofs> ed(nfields);
  for (int f = 0; f < nfields; ++f) {
    if (field_order[f] < 1 || field_order[f] > 2 || field_vdim[f] < 1) return FEM_ERR_ARG;
    ed[f] = entity_dofs(refshape, field_order[f]);
    const int nb = sh.nv * ed[f].vertex + (int)sh.edges.size() * ed[f].edge +
                   (sh.rdim == 3 ? (int)sh.faces.size() * ed[f].face : (sh.rdim == 2 ? ed[f].face : 0)) + ed[f].volume;
    ndpc += nb * field_vdim[f];
  }
  *ndofs_per_cell = ndpc;
  if (!cell_dofs) return FEM_OK;  // size query

  std::vector<std::vector<int64_t>> vdict(nfields, std::vector<int64_t>(nnodes + 1, 0));
  std::vector<std::unordered_map<std::pair<int64_t, int64_t>, int64_t, PairHash>> edict(nfields);
  std::vector<std::unordered_map<TripleKey, int64_t, TripleHash>> fdict(nfields);
  for (int f = 0; f < nfields; ++f) {
    if (ed[f].edge) edict[f].reserve((size_t)(2 * ncells));
    if (ed[f].face) fdict[f].reserve((size_t)(sh.rdim == 3 ? 2 * ncells : ncells));
  }
  int64_t nextdof = 1;
  int64_t pos = 0;
  for (int64_t ci = 0; ci < ncells; ++ci) {
    const int64_t* cn = cell_nodes + ci * npc;
    for (int f = 0; f < nfields; ++f) {
      const int nc = field_vdim[f];
      // vertices
      for (int v = 0; v < sh.nv; ++v) {
        const int64_t node = cn[v];
        if (node < 1 || node > nnodes) return FEM_ERR_ARG;
        int64_t& first = vdict[f][node];
        if (first > 0) {
          for (int d = 0; d < nc; ++d) cell_dofs[pos++] = first + d;
        } else {
          first = nextdof;
          for (int d = 0; d < nc; ++d) cell_dofs[pos++] = nextdof++;
        }
      }
      // edges (one interior dof for order 2, so orientation does not matter)
      if (ed[f].edge) {
        for (const auto& e : sh.edges) {
          int64_t a = cn[e.first], b = cn[e.second];
          if (a > b) std::swap(a, b);
          auto it = edict[f].find({a, b});
          int64_t first;
          if (it != edict[f].end()) {
            first = it->second;
          } else {
            first = nextdof;
            edict[f].emplace(std::make_pair(a, b), nextdof);
            nextdof += nc;
          }
          for (int d = 0; d < nc; ++d) cell_dofs[pos++] = first + d;
        }
      }
      // faces: key = three smallest node ids, sorted (sortface_fast)
      if (ed[f].face) {
        for (const auto& fc : sh.faces) {
          int64_t k[4];
          for (int i = 0; i < 4; ++i) k[i] = cn[fc[i]];
          std::sort(k, k + 4);
          const TripleKey key{k[0], k[1], k[2]};
          auto it = fdict[f].find(key);
          int64_t first;
          if (it != fdict[f].end()) {
            first = it->second;
          } else {
            first = nextdof;
            fdict[f].emplace(key, nextdof);
            nextdof += nc;
          }
          for (int d = 0; d < nc; ++d) cell_dofs[pos++] = first + d;
        }
      }
      for (int v = 0; v < ed[f].volume; ++v)
        for (int d = 0; d < nc; ++d) cell_dofs[pos++] = nextdof++;
    }
  }
  *ndofs = nextdof - 1;
  return FEM_OK;
}

int32_t fem_setup_pattern(int64_t ncells, int32_t ndpc, const int64_t* cell_dofs, int64_t ndofs, int64_t* colptr,
                          int64_t* rowval, int64_t capacity, int64_t* nnz) {
  if (!cell_dofs || !colptr || !nnz || ncells <= 0 || ndpc <= 0 || ndofs <= 0) return FEM_ERR_ARG;
  // dof -> cells (CSR)
  std::vector<int64_t> dptr(ndofs + 2, 0);
  for (int64_t i = 0; i < ncells * ndpc; ++i) {
    const int64_t d = cell_dofs[i];
    if (d < 1 || d > ndofs) return FEM_ERR_ARG;
    dptr[d + 1]++;
  }
  for (int64_t d = 1; d <= ndofs + 1; ++d) dptr[d] += dptr[d - 1];
  std::vector<int64_t> dcells(dptr[ndofs + 1]);
  {
    std::vector<int64_t> cur(dptr.begin(), dptr.end() - 1);
    for (int64_t c = 0; c < ncells; ++c)
      for (int a = 0; a < ndpc; ++a) dcells[cur[cell_dofs[c * ndpc + a]]++] = c;
  }
  std::vector<int64_t> counts(ndofs + 1, 0);
  const bool fill = rowval != nullptr;
  if (fill && capacity < colptr[ndofs] - 1) return FEM_ERR_ARG;
  std::atomic<int> bad{0};
  parallel_for(ndofs, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> buf;
    for (int64_t j = lo; j < hi; ++j) {
      const int64_t d = j + 1;
      buf.clear();
      buf.push_back(d);
      for (int64_t k = dptr[d]; k < dptr[d + 1]; ++k) {
        const int64_t c = dcells[k];
        buf.insert(buf.end(), cell_dofs + c * ndpc, cell_dofs + (c + 1) * ndpc);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      if (!fill) {
        counts[d] = (int64_t)buf.size();
      } else {
        const int64_t p0 = colptr[j] - 1;
        if (colptr[j + 1] - colptr[j] != (int64_t)buf.size()) { bad = 1; continue; }
        std::memcpy(rowval + p0, buf.data(), sizeof(int64_t) * buf.size());
      }
    }
  });
  if (bad) return FEM_ERR_ARG;
  if (!fill) {
    colptr[0] = 1;
    for (int64_t j = 0; j < ndofs; ++j) colptr[j + 1] = colptr[j] + counts[j + 1];
  }
  *nnz = colptr[ndofs] - 1;
  return FEM_OK;
}

}  // extern "C"

namespace {

// Neighbours (cells sharing >= 1 node), sorted unique, excluding the cell (coloring.jl:47-103).
void incidence(int64_t ncells, int npc, const int64_t* cn, int64_t nnodes, std::vector<int64_t>& ptr,
               std::vector<int32_t>& nb) {
  std::vector<int64_t> nptr(nnodes + 2, 0);
  for (int64_t i = 0; i < ncells * npc; ++i) nptr[cn[i] + 1]++;
  for (int64_t v = 1; v <= nnodes + 1; ++v) nptr[v] += nptr[v - 1];
  std::vector<int32_t> ncell(nptr[nnodes + 1]);
  {
    std::vector<int64_t> cur(nptr.begin(), nptr.end() - 1);
    for (int64_t c = 0; c < ncells; ++c)
      for (int a = 0; a < npc; ++a) ncell[cur[cn[c * npc + a]]++] = (int32_t)c;
  }
  std::vector<int64_t> cnt(ncells + 1, 0);
  std::vector<std::vector<int32_t>> chunks;
  const int nt = hw_threads();
  const int64_t chunk = (ncells + nt - 1) / nt;
  chunks.resize(nt);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([&, t] {
      std::vector<int32_t> cand;
      const int64_t lo = t * chunk, hi = std::min(ncells, lo + chunk);
      for (int64_t c = lo; c < hi; ++c) {
        cand.clear();
        for (int a = 0; a < npc; ++a) {
          const int64_t v = cn[c * npc + a];
          for (int64_t r = nptr[v]; r < nptr[v + 1]; ++r)
            if (ncell[r] != c) cand.push_back(ncell[r]);
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        cnt[c + 1] = (int64_t)cand.size();
        chunks[t].insert(chunks[t].end(), cand.begin(), cand.end());
      }
    });
  }
  for (auto& x : th) x.join();
  ptr.assign(ncells + 1, 0);
  for (int64_t c = 0; c < ncells; ++c) ptr[c + 1] = ptr[c] + cnt[c + 1];
  nb.clear();
  nb.reserve(ptr[ncells]);
  for (auto& ch : chunks) nb.insert(nb.end(), ch.begin(), ch.end());
}

// _greedy_coloring! (coloring.jl:108-145) over `cells` with membership test.
template <class M>
std::vector<std::vector<int32_t>> greedy(const std::vector<int32_t>& cells, const std::vector<int64_t>& ptr,
                                         const std::vector<int32_t>& nb, std::vector<int32_t>& color, M is_member) {
  std::vector<std::vector<int32_t>> fin;
  std::vector<char> occ;
  int total = 0;
  for (int32_t c : cells) {
    std::fill(occ.begin(), occ.begin() + total, 0);
    for (int64_t r = ptr[c]; r < ptr[c + 1]; ++r) {
      const int32_t o = nb[r];
      if (!is_member(o)) continue;
      const int32_t col = color[o];
      if (col != 0) occ[col - 1] = 1;
    }
    int free_c = 0;
    for (int k = 1; k <= total; ++k)
      if (!occ[k - 1]) { free_c = k; break; }
    if (free_c == 0) {
      ++total;
      free_c = total;
      if ((int)occ.size() < total) occ.push_back(0);
      fin.emplace_back();
    }
    color[c] = free_c;
    fin[free_c - 1].push_back(c);
  }
  return fin;
}

}  // namespace

extern "C" int32_t fem_setup_coloring(int64_t ncells, int32_t npc, const int64_t* cell_nodes_1based, int64_t nnodes,
                                      int32_t alg, int32_t* cell_color, int32_t* ncolors) {
  if (!cell_nodes_1based || !cell_color || !ncolors || ncells <= 0 || npc <= 0) return FEM_ERR_ARG;
  std::vector<int64_t> cn(ncells * (int64_t)npc);
  for (int64_t i = 0; i < ncells * npc; ++i) {
    if (cell_nodes_1based[i] < 1 || cell_nodes_1based[i] > nnodes) return FEM_ERR_ARG;
    cn[i] = cell_nodes_1based[i] - 1;
  }
  std::vector<int64_t> ptr;
  std::vector<int32_t> nb;
  incidence(ncells, npc, cn.data(), nnodes, ptr, nb);
  std::vector<int32_t> color(ncells, 0);
  std::vector<std::vector<int32_t>> fin;
  if (alg == FEM_COLORING_GREEDY) {
    std::vector<int32_t> cells(ncells);
    for (int64_t c = 0; c < ncells; ++c) cells[c] = (int32_t)c;
    fin = greedy(cells, ptr, nb, color, [](int32_t) { return true; });
  } else {
    if (ncells == 1) {
      fin = {{0}};
    } else {
      // 1. partitioning into BFS zones (coloring.jl:184-216)
      std::vector<int32_t> zone_of(ncells, 0);
      std::vector<std::vector<int32_t>> zones;
      int64_t visited = 0, seed_idx = 0;
      while (visited < ncells) {
        while (zone_of[seed_idx] != 0) ++seed_idx;
        zones.push_back({(int32_t)seed_idx});
        zone_of[seed_idx] = (int32_t)zones.size();
        ++visited;
        while (true) {
          std::vector<int32_t> s;
          const int32_t Z = (int32_t)zones.size() + 1;
          for (int32_t c : zones.back())
            for (int64_t r = ptr[c]; r < ptr[c + 1]; ++r) {
              const int32_t o = nb[r];
              if (zone_of[o] == 0) {
                zone_of[o] = Z;
                s.push_back(o);
              }
            }
          if (s.empty()) break;
          visited += (int64_t)s.size();
          zones.push_back(std::move(s));
        }
      }
      // 2. greedy coloring per zone (zones only see their own cells)
      std::vector<std::vector<std::vector<int32_t>>> zc(zones.size());
      for (size_t zi = 0; zi < zones.size(); ++zi) {
        const int32_t Z = (int32_t)zi + 1;
        zc[zi] = greedy(zones[zi], ptr, nb, color, [&](int32_t o) { return zone_of[o] == Z; });
      }
      // 3. gathering (coloring.jl:230-260)
      auto findmax_par = [&](bool odd, int64_t& best, size_t& arg) {
        best = INT64_MIN; arg = 0;
        bool first = true;
        for (size_t x = 1; x <= zc.size(); ++x) {
          const int64_t v = ((x % 2 == 1) == odd) ? (int64_t)zc[x - 1].size() : INT64_MIN;
          if (first || v > best) { best = v; arg = x; first = false; }
        }
      };
      int64_t Nodd, Neven;
      size_t Zodd, Zeven;
      findmax_par(true, Nodd, Zodd);
      findmax_par(false, Neven, Zeven);
      const int64_t N = Nodd + Neven;
      for (auto& v : zc[Zodd - 1]) fin.push_back(v);
      for (auto& v : zc[Zeven - 1]) fin.push_back(v);
      std::vector<int64_t> sizes(fin.size());
      for (size_t i = 0; i < fin.size(); ++i) sizes[i] = (int64_t)fin[i].size();
      for (size_t Z = 1; Z <= zc.size(); ++Z) {
        if (Z == Zodd || Z == Zeven) continue;
        auto& vecs = zc[Z - 1];
        const bool odd = Z % 2 == 1;
        std::vector<int64_t> used;
        std::vector<size_t> order(vecs.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return vecs[a].size() > vecs[b].size(); });
        for (size_t lc : order) {
          int64_t best = INT64_MAX, arg = 0;
          bool first = true;
          for (int64_t x = 1; x <= N; ++x) {
            const bool cond = odd ? (x <= Nodd) : (x > Nodd);
            const bool isused = std::find(used.begin(), used.end(), x) != used.end();
            const int64_t v = (cond && !isused) ? sizes[x - 1] : INT64_MAX;
            if (first || v < best) { best = v; arg = x; first = false; }
          }
          used.push_back(arg);
          auto& tgt = fin[arg - 1];
          tgt.insert(tgt.end(), vecs[lc].begin(), vecs[lc].end());
          sizes[arg - 1] = (int64_t)tgt.size();
        }
      }
    }
  }
  for (size_t k = 0; k < fin.size(); ++k)
    for (int32_t c : fin[k]) cell_color[c] = (int32_t)k + 1;
  *ncolors = (int32_t)fin.size();
  return FEM_OK;
}
