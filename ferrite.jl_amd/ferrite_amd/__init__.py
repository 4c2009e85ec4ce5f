"""ferrite_amd: Ferrite.jl's cell-loop assembly path on AMD MI355X (gfx950).

Python host interface over the C ABI of libferrite_mi355x.so (include/ferrite_mi355x.h).
Names follow Ferrite.jl 1.6.0; mutating Julia functions drop the `!`
(`add!` -> `add`, `close!` -> `close`, `apply!` -> `apply`).
"""
from ._lib import (FemArgumentError, FemConvergenceError, FemError, FemPatternError, LIB_PATH,
                   exported_symbols)
from .batch import BatchedPoisson, batch_sizes, make_batch
from .constraints import (AffineConstraint, ConformityConstraint, ConstraintHandler, Dirichlet, PeriodicDirichlet,
                          collect_periodic_facets)
from .device import (MI355XCSC, Context, ElasticityElement, PoissonElement, add_point_load, allocate_matrix,
                     allocate_vector, apply, apply_solution, apply_traction, apply_zero, compliance_gradient, simp_modulus, assemble_global, get_tuning, set_cell_parameters, set_tuning, solve,
                     spmv_bench, start_assemble, timings)
from .dofs import DofHandler, add, add_constraint_entries, close, create_coloring, create_sparsity_pattern
from .mesh import Grid, Hexahedron, NonConformingGrid, Quadrilateral, generate_grid
from .refel import (FacetQuadratureRule, Lagrange, QuadratureRule, RefHexahedron, RefLine, RefQuadrilateral,
                    VectorizedInterpolation)
from .values import CellValues, FacetValues
from .vtk import (VTKGridFile, write_cell_colors, write_cell_data, write_node_data, write_nodeset,
                  write_solution)

__all__ = [
    "FemArgumentError", "FemConvergenceError", "FemError", "FemPatternError", "LIB_PATH", "exported_symbols",
    "AffineConstraint", "ConformityConstraint", "ConstraintHandler", "Dirichlet", "PeriodicDirichlet",
    "collect_periodic_facets",
    "add_constraint_entries", "apply_solution", "MI355XCSC", "Context", "ElasticityElement", "PoissonElement",
    "allocate_matrix", "allocate_vector", "apply", "apply_traction", "add_point_load", "compliance_gradient",
    "simp_modulus", "apply_zero", "assemble_global", "set_cell_parameters",
    "solve", "spmv_bench", "set_tuning", "get_tuning", "start_assemble", "timings", "DofHandler", "add", "close", "create_coloring",
    "create_sparsity_pattern", "Grid", "Hexahedron", "NonConformingGrid", "Quadrilateral", "generate_grid",
    "FacetQuadratureRule",
    "Lagrange", "QuadratureRule", "RefHexahedron", "RefLine", "RefQuadrilateral", "VectorizedInterpolation",
    "CellValues", "FacetValues", "BatchedPoisson", "batch_sizes", "make_batch",
    "VTKGridFile", "write_solution", "write_node_data", "write_cell_data", "write_nodeset", "write_cell_colors",
]
