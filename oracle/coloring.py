"""Oracle: cell coloring (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates src/Grid/coloring.jl:47-103 (create_incidence_matrix), :108-160
(_greedy_coloring!, greedy_coloring), :171-263 (workstream_coloring) and
:308-317 (create_coloring) as literal loops.  Returns lists of 1-based cell ids.
"""


def create_incidence(grid):
    """For each cell (1-based), the sorted unique neighbours sharing >= 1 node (excluding itself)."""
    node_cells = {}
    for ci, cell in enumerate(grid.cells, start=1):
        for v in cell:
            node_cells.setdefault(v, []).append(ci)
    nbrs = [None]
    for ci, cell in enumerate(grid.cells, start=1):
        cand = []
        for v in cell:
            for c in node_cells[v]:
                if c != ci:
                    cand.append(c)
        nbrs.append(sorted(set(cand)))
    return nbrs


def _greedy(cell_colors, nbrs, cells, is_member):
    final = []
    total = 0
    for cellid in cells:
        occupied = [False] * (total + 1)
        for nb in nbrs[cellid]:
            if not is_member(nb):
                continue
            col = cell_colors[nb]
            if col != 0:
                occupied[col] = True
        free = 0
        for attempt in range(1, total + 1):
            if not occupied[attempt]:
                free = attempt
                break
        if free == 0:
            total += 1
            free = total
            final.append([])
        cell_colors[cellid] = free
        final[free - 1].append(cellid)
    return final


def greedy_coloring(grid):
    nbrs = create_incidence(grid)
    cell_colors = [0] * (grid.ncells + 1)
    return _greedy(cell_colors, nbrs, range(1, grid.ncells + 1), lambda c: True)


def workstream_coloring(grid):
    ncells = grid.ncells
    nbrs = create_incidence(grid)
    cellvec = list(range(1, ncells + 1))
    if ncells == 0:
        return []
    if ncells == 1:
        return [[1]]
    zone_of = [0] * (ncells + 1)
    zones = []
    n_visited = 0
    seed_idx = 0
    while n_visited < len(cellvec):
        while zone_of[cellvec[seed_idx]] != 0:
            seed_idx += 1
        seed = cellvec[seed_idx]
        zones.append([seed])
        zone_of[seed] = len(zones)
        n_visited += 1
        while True:
            s = []
            Z = len(zones) + 1
            for c in zones[-1]:
                for nb in nbrs[c]:
                    if zone_of[nb] == 0:
                        zone_of[nb] = Z
                        s.append(nb)
            if not s:
                break
            zones.append(s)
            n_visited += len(s)
    cell_colors = [0] * (ncells + 1)
    zone_colors = []
    for zi, zone in enumerate(zones, start=1):
        zone_colors.append(_greedy(cell_colors, nbrs, zone, lambda c, zi=zi: zone_of[c] == zi))

    def findmax_parity(odd):
        best, arg = None, None
        for x in range(1, len(zone_colors) + 1):
            v = len(zone_colors[x - 1]) if (x % 2 == 1) == odd else -(1 << 62)
            if best is None or v > best:
                best, arg = v, x
        return best, arg
    Nodd, Zodd = findmax_parity(True)
    Neven, Zeven = findmax_parity(False)
    N = Nodd + Neven
    final = [list(c) for c in zone_colors[Zodd - 1]] + [list(c) for c in zone_colors[Zeven - 1]]
    sizes = [len(c) for c in final]
    for Z in range(1, len(zone_colors) + 1):
        if Z == Zodd or Z == Zeven:
            continue
        vecs = zone_colors[Z - 1]
        odd = Z % 2 == 1
        used = set()
        order = sorted(range(len(vecs)), key=lambda k: -len(vecs[k]))  # stable, rev by length
        for lc in order:
            best, arg = None, None
            for x in range(1, N + 1):
                cond = (x <= Nodd) if odd else (x > Nodd)
                v = sizes[x - 1] if (cond and x not in used) else (1 << 62)
                if best is None or v < best:
                    best, arg = v, x
            used.add(arg)
            final[arg - 1].extend(vecs[lc])
            sizes[arg - 1] = len(final[arg - 1])
    for c in final:
        c.sort()
    return final


def create_coloring(grid, alg="workstream"):
    if alg == "workstream":
        return workstream_coloring(grid)
    if alg == "greedy":
        return greedy_coloring(grid)
    raise ValueError(alg)
