"""CPU oracle: a plain restatement of Ferrite.jl 1.6.0's cell-loop assembly path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline).  The product (``ferrite.jl_amd/``) never
imports, links or executes anything under ``oracle/``.

Each module follows the reference file:line it restates, written as literal
loops in the reference's own operation order, so that it can be read side by
side with the Julia source.  It is deliberately slow and meant for small grids
(C1 = 100x100 quads and below; 3D grids up to a few thousand cells).

Pinning (what checks each function against something other than itself):

=======================  ====================================================
function                 pinned by
=======================  ====================================================
grid.generate_grid       closed form node positions; test_grid_generators.jl
interpolations           Kronecker delta at reference nodes, partition of
                         unity, finite-difference gradients
quadrature               closed-form Gauss points/weights, polynomial
                         exactness (test_quadrules.jl:22-39)
dofs.distribute_dofs     celldofs KATs from test_dofs.jl:60-127
sparsity.pattern         brute-force (dofs x dofs per cell) + diagonal set
coloring                 validity invariant (test_grid_dofhandler_vtk.jl
                         :719-760), greedy 4/8-colour structure
cellvalues.reinit        affine reproduction, volume, spatial coordinate
                         (test_cellvalues.jl:60-175)
elements.poisson         closed-form tensor-product 1D stiffness/mass
elements.elasticity      symmetry, rigid-body null space, patch test
assembler                heat_equation.jl:234 norm(u) KAT (with apply+solve)
constraints              test_constraints.jl KATs; heat_equation.jl:234
solve                    scipy SuperLU direct solve (library routine)
adjoint                  central finite differences; torch float64 autograd
=======================  ====================================================

Bit-level parity at the Tensors.jl boundary (det, inv, dual-number gradient)
is unpinned: those routines are not vendored in the reference tree; the
restatement uses the textbook cofactor formulas, which agree to a few ulp.
"""
