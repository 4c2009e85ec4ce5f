"""Setup on the device (close!(dh) and allocate_matrix(dh)'s pattern, fem_dofs_distribute /
fem_pattern_build) against the host restatement of the reference's algorithms (fem_setup.cpp,
bit-exact with the oracle in tests/test_host_setup.py): cell_dofs, colptr and rowval equal.

Cases: scalar and vector Lagrange Q1 / Q2 on quadrilaterals and hexahedra, a mesh whose cell
order and node numbering are shuffled (so first visits are not in lattice order), and a 40³
hex mesh.
"""
import numpy as np
import pytest

import ferrite_amd as fa

pytestmark = pytest.mark.gpu

CASES = [("Quadrilateral", (7, 5), 1, 1), ("Quadrilateral", (6, 4), 1, 2), ("Quadrilateral", (5, 4), 2, 1),
         ("Quadrilateral", (4, 3), 2, 2), ("Hexahedron", (6, 5, 4), 1, 1), ("Hexahedron", (4, 3, 3), 1, 3),
         ("Hexahedron", (3, 2, 2), 2, 1), ("Hexahedron", (3, 3, 2), 2, 3)]


def _ip(g, order, vdim):
    ip = fa.Lagrange(g.refshape, order)
    return ip ** vdim if vdim > 1 else ip


def _compare(g, order, vdim):
    dh_h = fa.DofHandler(g).add("u", _ip(g, order, vdim)).close()
    dh_d = fa.DofHandler(g).add("u", _ip(g, order, vdim)).close(device=0)
    assert dh_d.ndofs == dh_h.ndofs and dh_d.ndofs_per_cell == dh_h.ndofs_per_cell
    assert np.array_equal(dh_d.cell_dofs, dh_h.cell_dofs)
    cp, rv = fa.create_sparsity_pattern(dh_h)
    K = fa.allocate_matrix(fa.MI355XCSC, dh_d)
    assert np.array_equal(K.colptr, cp) and np.array_equal(K.rowval, rv)
    return dh_d, K


@pytest.mark.parametrize("case", CASES)
def test_device_setup_matches_host(case):
    ct, nel, order, vdim = case
    print("case device setup", case, flush=True)
    _compare(fa.generate_grid(ct, nel), order, vdim)


@pytest.mark.parametrize("order,vdim", [(1, 1), (2, 3)])
def test_device_setup_shuffled_mesh(order, vdim):
    """Cells in random order, nodes renumbered at random: first visits jump around the mesh."""
    print("case device setup shuffled", order, vdim, flush=True)
    g = fa.generate_grid("Hexahedron", (5, 4, 3))
    rng = np.random.default_rng(4)
    perm_c = rng.permutation(g.ncells)
    perm_n = rng.permutation(g.nnodes)           # new id of old node k is perm_n[k]
    inv = np.argsort(perm_n)
    g.cells = (perm_n[g.cells[perm_c] - 1] + 1).astype(g.cells.dtype)
    g.nodes = g.nodes[inv]
    _compare(g, order, vdim)


def test_device_setup_40_cubed_and_assembly():
    print("case device setup 40^3", flush=True)
    g = fa.generate_grid("Hexahedron", (40, 40, 40))
    dh, K = _compare(g, 1, 1)
    # the device-built system assembles and solves like the host-built one
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), fa.Lagrange("RefHexahedron", 1))
    f = fa.allocate_vector(K)
    fa.assemble_global(fa.start_assemble(K, f), cv, fa.create_coloring(g), fa.PoissonElement())
    dh_h = fa.DofHandler(g).add("u", fa.Lagrange("RefHexahedron", 1)).close()
    Kh = fa.allocate_matrix(fa.MI355XCSC, dh_h)
    fh = fa.allocate_vector(Kh)
    fa.assemble_global(fa.start_assemble(Kh, fh), cv, fa.create_coloring(g), fa.PoissonElement())
    assert np.array_equal(K.nzval(), Kh.nzval()) and np.array_equal(f.to_numpy(), fh.to_numpy())
