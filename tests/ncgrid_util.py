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
