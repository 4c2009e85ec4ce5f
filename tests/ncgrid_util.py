"""Test helper: non-conforming quadrilateral grids by splitting chosen cells into four (the
hanging nodes of Ferrite's AMR forests after one refinement, without the forest)."""
from collections import defaultdict

import numpy as np

import ferrite_amd as fa


def refine_quads(grid, refine):
    """Split the quadrilaterals `refine` (1-based ids, no two sharing an edge) of a conforming
    grid into four children in Ferrite's corner order; the midpoint of a split edge shared
    with an unsplit cell is a hanging node whose masters are the edge's end nodes."""
    refine = set(int(c) for c in refine)
    nodes = [tuple(x) for x in grid.nodes]
    edge_cells = defaultdict(list)
    for ci, c in enumerate(grid.cells, 1):
        for a in range(4):
            edge_cells[tuple(sorted((int(c[a]), int(c[(a + 1) % 4]))))].append(ci)

    def add_node(x):
        nodes.append(tuple(x))
        return len(nodes)
    mid, hanging, cells = {}, {}, []
    for ci, c in enumerate(grid.cells, 1):
        c = [int(v) for v in c]
        if ci not in refine:
            cells.append(c)
            continue
        X = grid.nodes[np.array(c) - 1]
        m = []
        for a in range(4):
            e = tuple(sorted((c[a], c[(a + 1) % 4])))
            if e not in mid:
                mid[e] = add_node((X[a] + X[(a + 1) % 4]) / 2)
                others = [o for o in edge_cells[e] if o != ci]
                if others and others[0] not in refine:
                    hanging[mid[e]] = list(e)
            m.append(mid[e])
        ctr = add_node(X.mean(axis=0))
        cells += [[c[0], m[0], ctr, m[3]], [m[0], c[1], m[1], ctr], [ctr, m[1], c[2], m[2]], [m[3], ctr, m[2], c[3]]]
    g = fa.NonConformingGrid("Quadrilateral", np.array(cells), np.array(nodes), hanging)
    lo, hi = g.nodes.min(axis=0), g.nodes.max(axis=0)
    for name, d, v in (("left", 0, lo[0]), ("right", 0, hi[0]), ("bottom", 1, lo[1]), ("top", 1, hi[1])):
        g.addfacetset(name, lambda x, d=d, v=v: abs(x[d] - v) < 1e-12)
    return g


def _hex_lid(ex, ey, ez):
    """Local corner of offset (ex, ey, ez) in Ferrite's hexahedron order."""
    return ez * 4 + (3 - ex if ey else ex)


def refine_hexes(grid, refine):
    """Split the hexahedra `refine` (1-based ids) of a conforming grid into eight children in
    Ferrite's corner order (child corner (i, j, k) of child (cx, cy, cz) at the parent's
    half-grid point (cx + i, cy + j, cz + k)).  An edge midpoint on an edge that an unsplit
    cell also has is a hanging node with the edge's ends as masters; a face centre on a face
    shared with an unsplit cell is one with the face's four corners as masters."""
    refine = set(int(c) for c in refine)
    nodes = [tuple(x) for x in grid.nodes]
    cells0 = [[int(v) for v in c] for c in grid.cells]
    edge_cells, face_cells = defaultdict(set), defaultdict(set)
    for ci, c in enumerate(cells0, 1):
        for ax in range(3):
            for side in (0, 1):
                fn = [c[_hex_lid(*p)] for p in _face_points(ax, side)]
                face_cells[frozenset(fn)].add(ci)
        for e in _hex_edges(c):
            edge_cells[e].add(ci)
    new_node = {}

    def node_at(key, x):
        if key not in new_node:
            nodes.append(tuple(x))
            new_node[key] = len(nodes)
        return new_node[key]
    hanging, cells = {}, []
    for ci, c in enumerate(cells0, 1):
        if ci not in refine:
            cells.append(c)
            continue
        X = grid.nodes[np.array(c) - 1]
        half = {}
        for a in range(3):
            for b in range(3):
                for d in range(3):
                    pts = [(ex, ey, ez) for ex in ({0, 1} if a == 1 else {a // 2})
                           for ey in ({0, 1} if b == 1 else {b // 2}) for ez in ({0, 1} if d == 1 else {d // 2})]
                    ids = [c[_hex_lid(*p)] for p in pts]
                    if len(ids) == 1:
                        half[a, b, d] = ids[0]
                        continue
                    key = ("e" if len(ids) == 2 else "f" if len(ids) == 4 else "c",) + \
                        (tuple(sorted(ids)) if len(ids) < 8 else (ci,))
                    x = X[[_hex_lid(*p) for p in pts]].mean(axis=0)
                    nid = node_at(key, x)
                    half[a, b, d] = nid
                    if len(ids) == 2 and any(o not in refine for o in edge_cells[tuple(sorted(ids))]):
                        hanging[nid] = sorted(ids)
                    elif len(ids) == 4 and any(o not in refine for o in face_cells[frozenset(ids)] - {ci}):
                        hanging[nid] = [c[_hex_lid(*p)] for p in pts]
        for cz in range(2):
            for cy in range(2):
                for cx in range(2):
                    cells.append([half[cx + ex, cy + ey, cz + ez] for ez in range(2) for (ex, ey) in
                                  ((0, 0), (1, 0), (1, 1), (0, 1))])
    g = fa.NonConformingGrid("Hexahedron", np.array(cells), np.array(nodes), hanging)
    lo, hi = g.nodes.min(axis=0), g.nodes.max(axis=0)
    names = (("left", 0, lo[0]), ("right", 0, hi[0]), ("front", 1, lo[1]), ("back", 1, hi[1]),
             ("bottom", 2, lo[2]), ("top", 2, hi[2]))
    for name, d, v in names:
        g.addfacetset(name, lambda x, d=d, v=v: abs(x[d] - v) < 1e-12)
    return g


def _face_points(ax, side):
    """The four corner offsets of the hexahedron face normal to axis `ax` at `side`."""
    pts = []
    for u in range(2):
        for w in range(2):
            p = [0, 0, 0]
            p[ax] = side
            others = [i for i in range(3) if i != ax]
            p[others[0]], p[others[1]] = u, w
            pts.append(tuple(p))
    return pts


def _hex_edges(c):
    """The 12 edges of a hexahedron (node ids), each as a sorted pair."""
    out = []
    for ax in range(3):
        others = [i for i in range(3) if i != ax]
        for u in range(2):
            for w in range(2):
                p0, p1 = [0, 0, 0], [0, 0, 0]
                p0[others[0]] = p1[others[0]] = u
                p0[others[1]] = p1[others[1]] = w
                p1[ax] = 1
                out.append(tuple(sorted((c[_hex_lid(*p0)], c[_hex_lid(*p1)]))))
    return out
