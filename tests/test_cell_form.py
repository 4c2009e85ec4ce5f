"""The premises of the solver's cell form of K (CPU, against the oracle).

The cell-form SpMV applies K as Σ_c k_c K¹ with K¹ reduced to one value per corner-offset class
(|Δx| + 2|Δy| + 4|Δz| of the two corners, Ferrite's corner order), and sums a row's 27 (9)
neighbours through sums of k over the cells they share.  These tests pin, with the oracle's
element routine and assembly:
  * a box cell's element matrix has that class structure (to rounding), a sheared one does not;
  * the oracle's assembled and constrained K of a box grid with random k_e equals the cell form
    entry by entry (the identity the device's FP64 check verifies before the form is used);
  * the regrouped row sum equals the per-cell definition for random data.
"""
import numpy as np

from oracle import constraints as OC
from oracle.fem import poisson_element
from oracle.problems import assemble_poisson, poisson_conductivity, setup_poisson

# corner offsets in Ferrite's local order (Quadrilateral: the first four)
CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _bits(a):
    ox, oy, oz = CORNER[a]
    return ox + 2 * oy + 4 * oz


def _classes(Ke):
    """f[m] = K¹ of corner 0 and the corner with offset bits m; max deviation from that form."""
    nb = Ke.shape[0]
    f = np.zeros(8)
    for b in range(nb):
        f[_bits(b)] = Ke[0, b]
    dev = max(abs(Ke[a, b] - f[_bits(a) ^ _bits(b)]) for a in range(nb) for b in range(nb))
    return f, dev


def _unit_element(ct, x):
    grid, cv, *_ = setup_poisson(ct, (1, 1, 1) if ct == "Hexahedron" else (1, 1))
    Ke, _ = poisson_element(cv, x, 1.0, 0.0)
    return Ke


def test_box_element_has_offset_classes():
    box = np.array([[0.2 + 0.3 * ox, -0.1 + 0.7 * oy, 0.5 + 1.1 * oz] for ox, oy, oz in CORNER])
    Ke = _unit_element("Hexahedron", box)
    f, dev = _classes(Ke)
    assert dev <= 1e-14 * np.abs(Ke).max()
    # the classes differ for a non-cubic box (eight distinct values, none lost to symmetry)
    assert len({round(v, 12) for v in f}) == 8
    rect = np.array([[0.3 * ox, 0.8 * oy] for ox, oy, _ in CORNER[:4]])
    Kq = _unit_element("Quadrilateral", rect)
    _, devq = _classes(Kq)
    assert devq <= 1e-14 * np.abs(Kq).max()


def test_sheared_element_breaks_offset_classes():
    """A parallelepiped sheared in x along z is congruent across a lattice, but its K¹ depends
    on the sign of the offset: the host's class check must reject it."""
    shear = np.array([[0.3 * ox + 0.2 * oz, 0.7 * oy, 1.1 * oz] for ox, oy, oz in CORNER])
    Ke = _unit_element("Hexahedron", shear)
    _, dev = _classes(Ke)
    assert dev > 1e-6 * np.abs(Ke).max()


def test_assembled_box_grid_equals_cell_form():
    nel, right = (4, 3, 2), (0.8, 0.9, 0.5)
    k = poisson_conductivity(nel, seed=21)
    grid, cv, dh, K, f, ch = setup_poisson("Hexahedron", nel, (0.0, 0.0, 0.0), right, bc_sets=["left", "right"])
    assemble_poisson(grid, cv, dh, K, f, k_cell=k)
    OC.apply(K, f, ch)
    A = K.to_scipy().toarray()
    fcls, dev = _classes(_unit_element("Hexahedron", grid.getcoordinates(1)))
    assert dev <= 1e-14 * np.abs(fcls).max()
    # the cell form, assembled: E[d_a, d_b] = Σ_c k_c f[bits(a) ^ bits(b)]
    E = np.zeros_like(A)
    for ci in range(1, grid.ncells + 1):
        d = np.asarray(dh.celldofs(ci)) - 1
        for a in range(8):
            for b in range(8):
                E[d[a], d[b]] += k[ci - 1] * fcls[_bits(a) ^ _bits(b)]
    pres = np.asarray(ch.prescribed_dofs) - 1
    free = np.setdiff1d(np.arange(A.shape[0]), pres)
    scale = np.abs(A).max()
    assert np.abs(A[np.ix_(free, free)] - E[np.ix_(free, free)]).max() <= 1e-13 * scale
    # apply! leaves only the diagonal on prescribed rows and columns
    off = A.copy()
    off[pres, pres] = 0.0
    assert np.abs(off[pres, :]).max() == 0.0 and np.abs(off[:, pres]).max() == 0.0
    # an entry changed by 1e-4 of itself leaves the form by far more than the solver's FP32 bar
    i, j = free[3], free[4]
    if E[i, j] != 0.0:
        s = 1.0 / np.sqrt(np.abs(np.diag(A)))
        assert abs(1e-4 * A[i, j]) * s[i] * s[j] > 2.0 ** -22


def _row_direct(z, kc, f, dim):
    """Σ_e k_e Σ_b f[b ^ e] z[offset b − e]: z indexed [dz + 1][dy + 1][dx + 1]."""
    nb = 1 << dim
    acc = 0.0
    for e in range(nb):
        ex, ey, ez = e & 1, (e >> 1) & 1, e >> 2
        sub = 0.0
        for b in range(nb):
            bx, by, bz = b & 1, (b >> 1) & 1, b >> 2
            sub += f[b ^ e] * z[bz - ez + 1, by - ey + 1, bx - ex + 1]
        acc += kc[e] * sub
    return acc


def _row_grouped(z, kc, f, dim):
    """The SpMV's grouping: each neighbour d once, with the sum of k over the cells holding it."""
    acc = 0.0
    for dz in ((-1, 0, 1) if dim == 3 else (0,)):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ks = 0.0
                for e in range(1 << dim):
                    ex, ey, ez = e & 1, (e >> 1) & 1, e >> 2
                    if all(0 <= c + o <= 1 for c, o in ((ex, dx), (ey, dy), (ez, dz))):
                        ks += kc[e]
                acc += f[abs(dx) + 2 * abs(dy) + 4 * abs(dz)] * ks * z[dz + 1, dy + 1, dx + 1]
    return acc


def test_grouped_row_sum_equals_per_cell_definition():
    rng = np.random.default_rng(5)
    for dim in (3, 2):
        for _ in range(20):
            z = rng.normal(size=(3, 3, 3))
            if dim == 2:
                z[0] = z[2] = 0.0
            kc = rng.uniform(0.2, 3.0, size=8)
            f = rng.normal(size=8)
            a, b = _row_direct(z, kc, f, dim), _row_grouped(z, kc, f, dim)
            assert abs(a - b) <= 1e-13 * (1.0 + abs(a))
