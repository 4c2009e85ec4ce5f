/*
 * ferrite_mi355x_setup.h -- native host-side setup exported by libferrite_mi355x.so.
 *
 * A Julia caller already has Ferrite's own DofHandler, allocate_matrix pattern and
 * create_coloring result and passes them to the upload entry points of
 * ferrite_mi355x.h.  Hosts without Julia (the Python host interface of this
 * package) compute the same arrays with these routines, which reproduce the
 * reference's algorithms so that the result is identical:
 *
 *   fem_setup_distribute_dofs  close!(dh)                 src/Dofs/DofHandler.jl:474-794
 *   fem_setup_pattern          allocate_matrix(dh) pattern src/Dofs/sparsity_pattern.jl:223-253, :951-1016
 *   fem_setup_coloring         create_coloring(grid; alg) src/Grid/coloring.jl:47-317
 *
 * All integer arrays are Int64 and 1-based, like Ferrite's.  No device work.
 */
#ifndef FERRITE_MI355X_SETUP_H
#define FERRITE_MI355X_SETUP_H

#include <stdint.h>

#include "ferrite_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

enum fem_refshape { FEM_REF_LINE = 1, FEM_REF_QUADRILATERAL = 2, FEM_REF_HEXAHEDRON = 3 };
enum fem_coloring_alg { FEM_COLORING_GREEDY = 0, FEM_COLORING_WORKSTREAM = 1 };

/* Dof distribution for one SubDofHandler over all cells; fields in add! order, each a
 * Lagrange{refshape, order}^vdim (order 1 or 2).  Call with cell_dofs = NULL to get
 * *ndofs_per_cell only; otherwise cell_dofs holds ncells * ndofs_per_cell entries. */
int32_t fem_setup_distribute_dofs(int32_t refshape, int64_t ncells, int32_t nodes_per_cell,
                                  const int64_t* cell_nodes, int64_t nnodes, int32_t nfields,
                                  const int32_t* field_order, const int32_t* field_vdim,
                                  int64_t* cell_dofs, int32_t* ndofs_per_cell, int64_t* ndofs);

/* Two calls: rowval = NULL fills colptr[ndofs+1] and *nnz; then rowval (capacity >= nnz)
 * with the same colptr fills the row indices (ascending within each column). */
int32_t fem_setup_pattern(int64_t ncells, int32_t ndofs_per_cell, const int64_t* cell_dofs, int64_t ndofs,
                          int64_t* colptr, int64_t* rowval, int64_t capacity, int64_t* nnz);

/* Colour of every cell (1-based) such that cells sharing a node differ; colour lists
 * are the cells of each colour in ascending order, as create_coloring returns them. */
int32_t fem_setup_coloring(int64_t ncells, int32_t nodes_per_cell, const int64_t* cell_nodes, int64_t nnodes,
                           int32_t alg, int32_t* cell_color, int32_t* ncolors);

#ifdef __cplusplus
}
#endif
#endif /* FERRITE_MI355X_SETUP_H */
