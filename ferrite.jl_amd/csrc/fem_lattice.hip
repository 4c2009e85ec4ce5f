// Tensor-product numbering of the dofs of a scalar Q1 field, found from the cell topology.
//
// The mixed solver stores its FP32 copy of Â without column indices when the dofs sit on a
// lattice: every dof gets integer coordinates (ix, iy, iz), every pattern entry couples
// lattice neighbours (|Δ| ≤ 1 per axis), and the operator becomes planes of values indexed
// by (row, neighbour slot).  The numbering comes from the cells alone — node coordinates
// play no part, so distorted but logically structured meshes qualify — by a breadth-first
// walk over cells that places each cell's nodes at origin + the local corner offsets of
// Ferrite's reference cell (Quadrilateral: (0,0) (1,0) (1,1) (0,1); Hexahedron: the same
// square at z = 0, then at z = 1).  The walk fails on the first inconsistency: a node put at
// two places, two nodes at one place, a disconnected mesh or a bounding box that is not
// exactly filled.  Whether the pattern's entries are lattice neighbours is checked on the
// device by the solver (its operator is built only then).
#include <algorithm>
#include <climits>
#include <vector>

#include "fem_internal.h"

namespace fem {

namespace {
constexpr int kCorner3[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
}  // namespace

bool find_lattice(const fem_ctx* c, Lattice* L) {
  *L = Lattice{};
  const int npc = c->npc;
  const int dim = (c->sdim == 3 && npc == 8) ? 3 : (c->sdim == 2 && npc == 4) ? 2 : 0;
  if (dim == 0 || c->ndpc != npc || c->vdim != 1 || c->n != c->nnodes || c->ncells <= 0) return false;
  const int64_t nn = c->nnodes, nc = c->ncells;
  const int32_t* cn = c->h_cell_nodes.data();
  const int32_t* cd = c->h_cell_dofs.data();
  if ((int64_t)c->h_cell_nodes.size() != nc * npc || (int64_t)c->h_cell_dofs.size() != nc * npc) return false;

  // node -> dof (one dof per node, the same in every cell)
  std::vector<int32_t> dof_of(nn, -1);
  for (int64_t e = 0; e < nc; ++e)
    for (int a = 0; a < npc; ++a) {
      const int32_t nd = cn[e * npc + a], d = cd[e * npc + a];
      if (nd < 0 || nd >= nn || d < 0 || d >= c->n) return false;
      if (dof_of[nd] < 0) dof_of[nd] = d;
      else if (dof_of[nd] != d) return false;
    }
  {
    std::vector<uint8_t> seen(c->n, 0);
    for (int64_t i = 0; i < nn; ++i) {
      if (dof_of[i] < 0 || seen[dof_of[i]]) return false;
      seen[dof_of[i]] = 1;
    }
  }

  // node -> cells (CSR)
  std::vector<int64_t> nptr(nn + 1, 0);
  for (int64_t k = 0; k < nc * npc; ++k) ++nptr[cn[k] + 1];
  for (int64_t i = 0; i < nn; ++i) nptr[i + 1] += nptr[i];
  std::vector<int32_t> ncell(nc * npc);
  {
    std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t e = 0; e < nc; ++e)
      for (int a = 0; a < npc; ++a) ncell[fill[cn[e * npc + a]]++] = (int32_t)e;
  }

  // breadth-first placement of cells
  std::vector<int32_t> px(nn, INT_MIN), py(nn, INT_MIN), pz(nn, INT_MIN);
  std::vector<uint8_t> visited(nc, 0);
  std::vector<int32_t> queue;
  queue.reserve(nc);
  auto place = [&](int64_t e, int32_t ox, int32_t oy, int32_t oz) {
    for (int a = 0; a < npc; ++a) {
      const int32_t nd = cn[e * npc + a];
      const int32_t x = ox + kCorner3[a][0], y = oy + kCorner3[a][1], z = oz + (dim == 3 ? kCorner3[a][2] : 0);
      if (px[nd] == INT_MIN) {
        px[nd] = x; py[nd] = y; pz[nd] = z;
      } else if (px[nd] != x || py[nd] != y || pz[nd] != z) {
        return false;
      }
    }
    return true;
  };
  visited[0] = 1;
  queue.push_back(0);
  if (!place(0, 0, 0, 0)) return false;
  for (size_t h = 0; h < queue.size(); ++h) {
    const int64_t e = queue[h];
    for (int a = 0; a < npc; ++a) {
      const int32_t nd = cn[e * npc + a];
      for (int64_t q = nptr[nd]; q < nptr[nd + 1]; ++q) {
        const int32_t e2 = ncell[q];
        if (visited[e2]) continue;
        int b = 0;
        while (b < npc && cn[(int64_t)e2 * npc + b] != nd) ++b;
        if (b == npc) return false;
        if (!place(e2, px[nd] - kCorner3[b][0], py[nd] - kCorner3[b][1],
                   pz[nd] - (dim == 3 ? kCorner3[b][2] : 0)))
          return false;
        visited[e2] = 1;
        queue.push_back(e2);
      }
    }
  }
  if ((int64_t)queue.size() != nc) return false;   // more than one connected piece

  int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int64_t i = 0; i < nn; ++i) {
    if (px[i] == INT_MIN) return false;             // a node in no cell
    lo[0] = std::min(lo[0], px[i]); hi[0] = std::max(hi[0], px[i]);
    lo[1] = std::min(lo[1], py[i]); hi[1] = std::max(hi[1], py[i]);
    lo[2] = std::min(lo[2], pz[i]); hi[2] = std::max(hi[2], pz[i]);
  }
  const int64_t nx = (int64_t)hi[0] - lo[0] + 1, ny = (int64_t)hi[1] - lo[1] + 1, nz = (int64_t)hi[2] - lo[2] + 1;
  if (nx * ny * nz != nn) return false;
  std::vector<int32_t> pos(c->n);
  std::vector<uint8_t> taken(nn, 0);
  for (int64_t i = 0; i < nn; ++i) {
    const int64_t p = (px[i] - lo[0]) + nx * ((py[i] - lo[1]) + ny * (int64_t)(pz[i] - lo[2]));
    if (taken[p]) return false;
    taken[p] = 1;
    pos[dof_of[i]] = (int32_t)p;
  }
  L->dim = dim;
  L->nx = nx; L->ny = ny; L->nz = nz;
  L->pos = std::move(pos);
  return true;
}

}  // namespace fem
