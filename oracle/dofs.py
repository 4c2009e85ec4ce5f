"""Oracle: DofHandler dof distribution (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates src/Dofs/DofHandler.jl:493-574 (__close!), :576-683
(_close_subdofhandler!), :685-713 (_distribute_dofs_for_cell!), :715-744
(add_vertex_dofs), :746-759 (get_or_create_dofs!), :761-794 (edge/face/volume),
:854-857 (sortedge), :1043-1059 (sortface) and permute_and_push! (:826-842)
for a single SubDofHandler covering all cells (the only case on this path).
Dofs are numbered on first visit, cells ascending, fields in add! order.
"""
from .refel import REF_EDGES, REF_FACES, base_interpolation, n_copies


def _sortedge(e):
    a, b = e
    return ((a, b), True) if a < b else ((b, a), False)


def _sortface(f):
    """sortface_fast: the three smallest node ids, sorted."""
    return tuple(sorted(f)[:3])


class DofHandler:
    def __init__(self, grid):
        self.grid = grid
        self.field_names = []
        self.field_ips = []
        self.closed = False

    def add(self, name, ip):
        self.field_names.append(name)
        self.field_ips.append(ip)
        return self

    def close(self):
        grid = self.grid
        shape = grid.refshape
        nf = len(self.field_names)
        vertexdicts = [dict() for _ in range(nf)]
        edgedicts = [dict() for _ in range(nf)]
        facedicts = [dict() for _ in range(nf)]
        nextdof = 1
        cell_dofs = []
        offsets = []
        self.field_offsets = []
        for ci in range(1, grid.ncells + 1):
            cell = grid.cells[ci - 1]
            offsets.append(len(cell_dofs))
            start = len(cell_dofs)
            for fi, ip in enumerate(self.field_ips):
                if ci == 1:
                    self.field_offsets.append(len(cell_dofs) - start)
                bip = base_interpolation(ip)
                nc = n_copies(ip)
                nv = [len(t) for t in bip.vertexdof_indices()]
                ne = [len(t) for t in bip.edgedof_interior_indices()]
                nfd = [len(t) for t in bip.facedof_interior_indices()]
                nvol = len(bip.volumedof_interior_indices())
                # vertices (add_vertex_dofs)
                for vi, vertex in enumerate(cell[:len(nv)]):
                    if nv[vi] == 0:
                        continue
                    first = vertexdicts[fi].get(vertex, 0)
                    if first > 0:
                        for lvi in range(nv[vi]):
                            for d in range(nc):
                                cell_dofs.append(first + lvi * nc + d)
                    else:
                        vertexdicts[fi][vertex] = nextdof
                        for _ in range(nv[vi] * nc):
                            cell_dofs.append(nextdof)
                            nextdof += 1
                # edges (add_edge_dofs)
                for ei, e in enumerate(REF_EDGES[shape]):
                    if ei >= len(ne) or ne[ei] == 0:
                        continue
                    key, regular = _sortedge((cell[e[0] - 1], cell[e[1] - 1]))
                    nextdof, dofs = _get_or_create(nextdof, ne[ei], nc, edgedicts[fi], key)
                    if not regular and bip.order > 2:  # adjust_during_distribution
                        dofs = dofs[::-1]
                    for dof in dofs:
                        for i in range(nc):
                            cell_dofs.append(dof + i)
                # faces (add_face_dofs)
                for fidx, f in enumerate(REF_FACES[shape]):
                    if fidx >= len(nfd) or nfd[fidx] == 0:
                        continue
                    key = _sortface(tuple(cell[v - 1] for v in f))
                    nextdof, dofs = _get_or_create(nextdof, nfd[fidx], nc, facedicts[fi], key)
                    for dof in dofs:
                        for i in range(nc):
                            cell_dofs.append(dof + i)
                # volume (add_volume_dofs)
                for _ in range(nvol * nc):
                    cell_dofs.append(nextdof)
                    nextdof += 1
            if ci == 1:
                self.ndofs_per_cell = len(cell_dofs) - start
        self.cell_dofs = cell_dofs
        self.cell_dofs_offset = offsets
        self.ndofs = nextdof - 1
        self.closed = True
        return self

    def celldofs(self, ci):
        o = self.cell_dofs_offset[ci - 1]
        return self.cell_dofs[o:o + self.ndofs_per_cell]

    def dof_range(self, name):
        fi = self.field_names.index(name)
        start = self.field_offsets[fi]
        return range(start, start + self.field_ips[fi].nbasefunctions)


def _get_or_create(nextdof, ndofs, nc, d, key):
    """get_or_create_dofs! (DofHandler.jl:746-759): a StepRange first:nc:(first+nc*ndofs-1)."""
    if key in d:
        first = d[key]
        return nextdof, list(range(first, first + nc * ndofs, nc))
    d[key] = nextdof
    dofs = list(range(nextdof, nextdof + nc * ndofs, nc))
    return nextdof + ndofs * nc, dofs
