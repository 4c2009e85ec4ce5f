/*
 * Oracle (TEST INFRASTRUCTURE ONLY): C restatement of Ferrite.jl's threaded
 * colored assembly loop (Q1 hex Poisson with fixed tables; scalar Poisson and
 * small-strain elasticity on any Lagrange interpolation from tabulated shape
 * functions), apply!, a Jacobi-PCG solve and the compliance adjoint, used as the
 * timed CPU baselines ("Ferrite-algorithm CPU restatement") and checked against
 * the Python oracle by the CPU tests.  Never linked into the product.
 *
 * Follows, per cell (docs/src/literate-howto/threaded_assembly.jl:232-265, with
 * OpenMP tasks over the cells of each colour, colours in sequence):
 *   reinit!(cellcache)            src/iterators.jl:72-88      gather coords, dofs
 *   reinit!(cellvalues)           src/FEValues/CellValues.jl:122-140
 *     J = Σ x_j ⊗ dMdξ_j          GeometryMapping.jl:124-130
 *     detJ > 0, detJdV = detJ w   CellValues.jl:110-116
 *     dNdx = dNdξ ⋅ J⁻¹           FunctionValues.jl:163, :186-192
 *   Ke += k (∇Nᵢ·∇Nⱼ) dΩ, fe += s Nᵢ dΩ   heat_equation.jl:143-164
 *   Ke += (∇ˢδuᵢ ⊡ C ⊡ ∇ˢuⱼ) dΩ           threaded_assembly.jl:105-120
 *   assemble!(A, dofs, Ke, fe)    src/assembler.jl:322-457: f[dofs] += fe; sort the
 *                                 dofs with a permutation; per sorted column merge-walk
 *                                 rowval against the sorted rows, skipping zero values
 * then apply!(K, f, ch)           src/Dofs/ConstraintHandler.jl:710-747 (g = 0)
 * and u = K \ f replaced by Jacobi-preconditioned CG (the GPU path's solver).
 * The compliance adjoint (dJ/dθ_e = −u_eᵀ ∂Ke/∂θ_e u_e and dJ/dX = −uᵀ ∂K/∂X u,
 * docs/src/literate-tutorials/topology_optimization.jl:320-383 for the SIMP element)
 * differentiates reinit! by hand: δJ = e_m ⊗ ∇_ξM_b, δ∇N_a = −(∇N_a)_m ∇M_b,
 * δdetJ = detJ (∇M_b)_m.
 * Indices are 0-based int32 here.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double now(void) { return omp_get_wtime(); }

/* Q1 tables at the 2x2x2 Gauss points (first coordinate fastest), hex local node order. */
static void q1hex_tables(double N[8][8], double dN[8][8][3], double w[8]) {
  const double g = 1.0 / sqrt(3.0);
  static const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                   sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  for (int q = 0; q < 8; ++q) {
    const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g, zeta = (q & 4) ? g : -g;
    w[q] = 1.0;
    for (int a = 0; a < 8; ++a) {
      const double fx = 1 + sx[a] * xi, fy = 1 + sy[a] * eta, fz = 1 + sz[a] * zeta;
      N[q][a] = fx * fy * fz / 8;
      dN[q][a][0] = sx[a] * fy * fz / 8;
      dN[q][a][1] = fx * sy[a] * fz / 8;
      dN[q][a][2] = fx * fy * sz[a] / 8;
    }
  }
}

/* insertion sort of the local dofs with the permutation (assembler.jl:513-576) */
static void sortperm8(const int32_t* dofs, int32_t* sorted, int* perm) {
  for (int i = 0; i < 8; ++i) { sorted[i] = dofs[i]; perm[i] = i; }
  for (int i = 1; i < 8; ++i) {
    int32_t v = sorted[i];
    int p = perm[i], j = i - 1;
    while (j >= 0 && sorted[j] > v) { sorted[j + 1] = sorted[j]; perm[j + 1] = perm[j]; --j; }
    sorted[j + 1] = v;
    perm[j + 1] = p;
  }
}

/* Returns 0 on success, -1 on det(J) <= 0, -2 on a missing pattern entry. */
int cb_assemble_q1hex(int64_t ncells, const double* xyz, const int32_t* cell_nodes, const int32_t* cell_dofs,
                      const double* kcell, double source, int64_t n, const int32_t* colptr, const int32_t* rowval,
                      int ncolors, const int64_t* color_ptr, const int32_t* color_cells, int nthreads,
                      double* nzval, double* f) {
  double N[8][8], dN[8][8][3], w[8];
  q1hex_tables(N, dN, w);
  const int64_t nnz = colptr[n];
  int status = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < nnz; ++p) nzval[p] = 0.0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) f[i] = 0.0;
  for (int col = 0; col < ncolors; ++col) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t s = color_ptr[col]; s < color_ptr[col + 1]; ++s) {
      const int32_t cell = color_cells[s];
      double x[8][3];
      for (int a = 0; a < 8; ++a)
        for (int d = 0; d < 3; ++d) x[a][d] = xyz[3 * (int64_t)cell_nodes[8 * (int64_t)cell + a] + d];
      const int32_t* dofs = cell_dofs + 8 * (int64_t)cell;
      double Ke[8][8] = {{0}}, fe[8] = {0};
      const double k = kcell[cell];
      for (int q = 0; q < 8; ++q) {
        double J[3][3] = {{0}};
        for (int a = 0; a < 8; ++a)
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] += x[a][i] * dN[q][a][j];
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[0][2] * J[2][1] - J[0][1] * J[2][2],
                     c02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        const double det = J[0][0] * c00 + J[1][0] * c01 + J[2][0] * c02;
        if (!(det > 0.0)) {
#pragma omp atomic write
          status = -1;
          continue;
        }
        const double id = 1.0 / det;
        double Ji[3][3] = {{c00 * id, c01 * id, c02 * id},
                           {(J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
                            (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
                           {(J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                            (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
        const double dV = det * w[q];
        double g[8][3];
        for (int b = 0; b < 8; ++b)
          for (int j = 0; j < 3; ++j) g[b][j] = dN[q][b][0] * Ji[0][j] + dN[q][b][1] * Ji[1][j] + dN[q][b][2] * Ji[2][j];
        for (int a = 0; a < 8; ++a) {
          fe[a] += source * N[q][a] * dV;
          for (int b = 0; b < 8; ++b) Ke[a][b] += k * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]) * dV;
        }
      }
      for (int a = 0; a < 8; ++a) f[dofs[a]] += fe[a];
      int32_t sd[8];
      int perm[8];
      sortperm8(dofs, sd, perm);
      for (int cj = 0; cj < 8; ++cj) {
        const int32_t kcol = sd[cj];
        const int kec = perm[cj];
        int32_t R = colptr[kcol];
        const int32_t Rend = colptr[kcol + 1];
        int ri = 0;
        while (R < Rend && ri < 8) {  /* merge walk (assembler.jl:431-455) */
          const int32_t krow = rowval[R];
          if (krow == sd[ri]) {
            const double v = Ke[perm[ri]][kec];
            if (v != 0.0) nzval[R] += v;
            ++ri;
            ++R;
          } else if (krow < sd[ri]) {
            ++R;
          } else {
            if (Ke[perm[ri]][kec] != 0.0) {
#pragma omp atomic write
              status = -2;
            }
            ++ri;
          }
        }
        for (; ri < 8; ++ri)
          if (Ke[perm[ri]][kec] != 0.0) {
#pragma omp atomic write
            status = -2;
          }
      }
    }
  }
  return status;
}

/* apply!(K, f, ch) with homogeneous values (ConstraintHandler.jl:710-747). Returns meandiag. */
double cb_apply_zero_dirichlet(int64_t n, const int32_t* colptr, const int32_t* rowval, double* nzval, double* f,
                               int64_t np, const int32_t* prescribed, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double z = 0.0;
  for (int64_t i = 0; i < n; ++i) {  /* meandiag, sequential like the reference */
    for (int32_t p = colptr[i]; p < colptr[i + 1]; ++p)
      if (rowval[p] == i) { z += fabs(nzval[p]); break; }
  }
  const double m = z / (double)n;
  unsigned char* isc = (unsigned char*)calloc((size_t)n, 1);
  for (int64_t i = 0; i < np; ++i) isc[prescribed[i]] = 1;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j)
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p)
      if (isc[j] || isc[rowval[p]]) nzval[p] = 0.0;
  for (int64_t i = 0; i < np; ++i) {
    const int32_t d = prescribed[i];
    for (int32_t p = colptr[d]; p < colptr[d + 1]; ++p)
      if (rowval[p] == d) nzval[p] = m;
    f[d] = 0.0 * m;
  }
  free(isc);
  return m;
}

/* Jacobi-PCG on the symmetric K (CSC columns used as rows).  Returns iterations. */
int cb_pcg(int64_t n, const int32_t* colptr, const int32_t* rowval, const double* nzval, const double* b, double* x,
           double rtol, int maxit, int nthreads, double* relres) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
         *q = malloc(sizeof(double) * n), *dinv = malloc(sizeof(double) * n);
  double bb = 0.0, rz = 0.0;
#pragma omp parallel for reduction(+ : bb, rz) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t k = colptr[i]; k < colptr[i + 1]; ++k)
      if (rowval[k] == i) d = nzval[k];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    x[i] = 0.0;
    r[i] = b[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
    bb += b[i] * b[i];
    rz += r[i] * z[i];
  }
  const double tol2 = rtol * rtol * bb;
  double rr = bb;
  int it = 0;
  while (it < maxit && rr > tol2) {
    double pq = 0.0;
#pragma omp parallel for reduction(+ : pq) schedule(static)
    for (int64_t j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) acc += nzval[k] * p[rowval[k]];
      q[j] = acc;
      pq += p[j] * acc;
    }
    const double alpha = rz / pq;
    double rzn = 0.0, rrn = 0.0;
#pragma omp parallel for reduction(+ : rzn, rrn) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
      rzn += r[i] * z[i];
      rrn += r[i] * r[i];
    }
    const double beta = rzn / rz;
    rz = rzn;
    rr = rrn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  *relres = sqrt(rr / bb);
  free(r); free(z); free(p); free(q); free(dinv);
  return it;
}

double cb_wtime(void) { return now(); }
int cb_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------------------------
 * Generic cell loop: any Lagrange interpolation given as tables at the quadrature points.
 *   N[q*nb + i], dN[(q*nb + i)*dim + k]   function base (scalar), dN = ∂N/∂ξ_k
 *   dM[(q*ng + a)*dim + k]                geometric base (Lagrange Q1 of the cell)
 *   w[q]                                  quadrature weights
 * Sizes: dim <= 3, ng <= 8, nb <= 27, nq <= 27, local dofs nd = nb (scalar) or nb*dim
 * (vector, components interleaved: local dof I = dim*i + c).
 * ------------------------------------------------------------------------------------------ */
enum { CB_MAXD = 3, CB_MAXG = 8, CB_MAXB = 27, CB_MAXQ = 27, CB_MAXND = 81 };

/* insertion sort of nd local dofs with the permutation (assembler.jl:513-576) */
static void sortperm_n(int nd, const int32_t* dofs, int32_t* sorted, int* perm) {
  for (int i = 0; i < nd; ++i) { sorted[i] = dofs[i]; perm[i] = i; }
  for (int i = 1; i < nd; ++i) {
    int32_t v = sorted[i];
    int p = perm[i], j = i - 1;
    while (j >= 0 && sorted[j] > v) { sorted[j + 1] = sorted[j]; perm[j + 1] = perm[j]; --j; }
    sorted[j + 1] = v;
    perm[j + 1] = p;
  }
}

/* assemble!(A, dofs, Ke): per sorted column, merge walk of rowval against the sorted rows
 * (assembler.jl:420-457).  Ke is nd x nd, row-major.  Returns 0, or -2 on a missing entry. */
static int scatter_n(int nd, const int32_t* dofs, const double* Ke, const int32_t* colptr, const int32_t* rowval,
                     double* nzval) {
  int32_t sd[CB_MAXND];
  int perm[CB_MAXND];
  int st = 0;
  sortperm_n(nd, dofs, sd, perm);
  for (int cj = 0; cj < nd; ++cj) {
    const int32_t kcol = sd[cj];
    const int kec = perm[cj];
    int32_t R = colptr[kcol];
    const int32_t Rend = colptr[kcol + 1];
    int ri = 0;
    while (R < Rend && ri < nd) {
      const int32_t krow = rowval[R];
      if (krow == sd[ri]) {
        const double v = Ke[perm[ri] * nd + kec];
        if (v != 0.0) nzval[R] += v;
        ++ri;
        ++R;
      } else if (krow < sd[ri]) {
        ++R;
      } else {
        if (Ke[perm[ri] * nd + kec] != 0.0) st = -2;
        ++ri;
      }
    }
    for (; ri < nd; ++ri)
      if (Ke[perm[ri] * nd + kec] != 0.0) st = -2;
  }
  return st;
}

/* reinit! at one quadrature point: J = Σ_a x_a ⊗ ∇_ξM_a, detJ, J⁻¹, and ∇N_i = ∇_ξN_i · J⁻¹
 * for the nb function base functions.  Returns detJ (the caller rejects detJ <= 0). */
static double reinit_q(int dim, int nb, int ng, const double* x, const double* dNq, const double* dMq,
                       double Ji[CB_MAXD][CB_MAXD], double g[CB_MAXB][CB_MAXD]) {
  double J[CB_MAXD][CB_MAXD] = {{0}};
  for (int a = 0; a < ng; ++a)
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j) J[i][j] += x[a * dim + i] * dMq[a * dim + j];
  double det;
  if (dim == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
    Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[0][2] * J[2][1] - J[0][1] * J[2][2],
                 c02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    det = J[0][0] * c00 + J[1][0] * c01 + J[2][0] * c02;
    const double id = 1.0 / det;
    Ji[0][0] = c00 * id; Ji[0][1] = c01 * id; Ji[0][2] = c02 * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  }
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < dim; ++j) {
      double acc = 0.0;
      for (int k = 0; k < dim; ++k) acc += dNq[b * dim + k] * Ji[k][j];
      g[b][j] = acc;
    }
  return det;
}

static void gather_x(int dim, int ng, const double* xyz, const int32_t* nodes, double* x) {
  for (int a = 0; a < ng; ++a)
    for (int d = 0; d < dim; ++d) x[a * dim + d] = xyz[(int64_t)dim * nodes[a] + d];
}

static int check_sizes(int dim, int nb, int ng, int nq, int vdim) {
  return dim >= 2 && dim <= CB_MAXD && nb >= 1 && nb <= CB_MAXB && ng >= 1 && ng <= CB_MAXG && nq >= 1 &&
         nq <= CB_MAXQ && nb * vdim <= CB_MAXND;
}

/* Scalar Poisson −∇·(k∇u) = s (heat_equation.jl:143-164) on any Lagrange base.  Zeroes K and f
 * first (start_assemble).  Returns 0, -1 on det(J) <= 0, -2 on a missing pattern entry, -3 on
 * sizes beyond the tables' limits. */
int cb_assemble_poisson(int dim, int nb, int ng, int nq, const double* N, const double* dN, const double* dM,
                        const double* w, int64_t ncells, const double* xyz, const int32_t* cell_nodes,
                        const int32_t* cell_dofs, const double* kcell, double source, int64_t n,
                        const int32_t* colptr, const int32_t* rowval, int ncolors, const int64_t* color_ptr,
                        const int32_t* color_cells, int nthreads, double* nzval, double* f) {
  if (!check_sizes(dim, nb, ng, nq, 1)) return -3;
  const int64_t nnz = colptr[n];
  int status = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < nnz; ++p) nzval[p] = 0.0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) f[i] = 0.0;
  for (int col = 0; col < ncolors; ++col) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = color_ptr[col]; s < color_ptr[col + 1]; ++s) {
      const int32_t cell = color_cells[s];
      double x[CB_MAXG * CB_MAXD], Ke[CB_MAXB * CB_MAXB], fe[CB_MAXB];
      double Ji[CB_MAXD][CB_MAXD], g[CB_MAXB][CB_MAXD];
      gather_x(dim, ng, xyz, cell_nodes + (int64_t)ng * cell, x);
      const int32_t* dofs = cell_dofs + (int64_t)nb * cell;
      memset(Ke, 0, sizeof(double) * nb * nb);
      memset(fe, 0, sizeof(double) * nb);
      const double k = kcell[cell];
      int bad = 0;
      for (int q = 0; q < nq; ++q) {
        const double det = reinit_q(dim, nb, ng, x, dN + (int64_t)q * nb * dim, dM + (int64_t)q * ng * dim, Ji, g);
        if (!(det > 0.0)) { bad = 1; break; }
        const double dV = det * w[q];
        for (int i = 0; i < nb; ++i) {
          fe[i] += source * N[q * nb + i] * dV;
          for (int j = 0; j < nb; ++j) {
            double dot = 0.0;
            for (int d = 0; d < dim; ++d) dot += g[i][d] * g[j][d];
            Ke[i * nb + j] += k * dot * dV;
          }
        }
      }
      if (bad) {
#pragma omp atomic write
        status = -1;
        continue;
      }
      for (int i = 0; i < nb; ++i) f[dofs[i]] += fe[i];
      if (scatter_n(nb, dofs, Ke, colptr, rowval, nzval)) {
#pragma omp atomic write
        status = -2;
      }
    }
  }
  return status;
}

/* Small-strain linear elasticity with the isotropic tensor C = λ I⊗I + 2μ I^sym (plane strain
 * for dim 2) and per-cell Young's modulus E[cell] (threaded_assembly.jl:105-132): Ke[I,J] +=
 * (ε_I ⊡ C ⊡ ε_J) dΩ with ε_I = sym(e_c ⊗ ∇N_i), I = dim*i + c.  As with Tensors.jl's
 * SymmetricTensor, the double contraction runs over the nv = 6 (3 in 2D) independent
 * components, off-diagonal ones counted twice; σ_J = C ⊡ ε_J is formed once per quadrature
 * point.  The dim-specialised body lets the compiler unroll the component loops and vectorise
 * over J.  Zeroes K first.  Same return codes as cb_assemble_poisson. */
static inline int elast_cell(const int dim, int nb, int ng, int nq, const double* dN, const double* dM,
                             const double* w, const double* x, double lam, double mu, double* Ke, double* ev,
                             double* sv) {
  const int nd = nb * dim, nv = dim == 3 ? 6 : 3;
  static const int vk3[6] = {0, 1, 2, 0, 0, 1}, vl3[6] = {0, 1, 2, 1, 2, 2};
  static const int vk2[3] = {0, 1, 0}, vl2[3] = {0, 1, 1};
  const int* vk = dim == 3 ? vk3 : vk2;
  const int* vl = dim == 3 ? vl3 : vl2;
  double Ji[CB_MAXD][CB_MAXD], g[CB_MAXB][CB_MAXD];
  memset(Ke, 0, sizeof(double) * nd * nd);
  for (int q = 0; q < nq; ++q) {
    const double det = reinit_q(dim, nb, ng, x, dN + (int64_t)q * nb * dim, dM + (int64_t)q * ng * dim, Ji, g);
    if (!(det > 0.0)) return -1;
    const double dV = det * w[q];
    /* ev[v*nd + I] = ε_I component v; sv[v*nd + J] = (C ⊡ ε_J) component v, times 2 off the
     * diagonal, times dV */
    for (int I = 0; I < nd; ++I) {
      const int i = I / dim, c = I % dim;
      double tr = g[i][c];
      for (int v = 0; v < nv; ++v) {
        const int k = vk[v], l = vl[v];
        const double e = 0.5 * ((k == c ? g[i][l] : 0.0) + (l == c ? g[i][k] : 0.0));
        ev[v * nd + I] = e;
        sv[v * nd + I] = (k == l ? 2.0 * mu * e + lam * tr : 2.0 * (2.0 * mu * e)) * dV;
      }
    }
    for (int I = 0; I < nd; ++I) {
      double* row = Ke + I * nd;
      for (int v = 0; v < nv; ++v) {
        const double e = ev[v * nd + I];
        const double* s = sv + v * nd;
        for (int J = 0; J < nd; ++J) row[J] += e * s[J];
      }
    }
  }
  return 0;
}

int cb_assemble_elasticity(int dim, int nb, int ng, int nq, const double* dN, const double* dM, const double* w,
                           int64_t ncells, const double* xyz, const int32_t* cell_nodes, const int32_t* cell_dofs,
                           const double* Ecell, double nu, int64_t n, const int32_t* colptr,
                           const int32_t* rowval, int ncolors, const int64_t* color_ptr, const int32_t* color_cells,
                           int nthreads, double* nzval) {
  if (!check_sizes(dim, nb, ng, nq, dim)) return -3;
  const int nd = nb * dim;
  const int64_t nnz = colptr[n];
  const double lam1 = nu / ((1 + nu) * (1 - 2 * nu)), mu1 = 1.0 / (2 * (1 + nu));
  int status = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < nnz; ++p) nzval[p] = 0.0;
#pragma omp parallel
  {
    double* Ke = (double*)malloc(sizeof(double) * CB_MAXND * CB_MAXND);
    double* ev = (double*)malloc(sizeof(double) * 6 * CB_MAXND);
    double* sv = (double*)malloc(sizeof(double) * 6 * CB_MAXND);
    for (int col = 0; col < ncolors; ++col) {
#pragma omp for schedule(dynamic, 16)
      for (int64_t s = color_ptr[col]; s < color_ptr[col + 1]; ++s) {
        const int32_t cell = color_cells[s];
        double x[CB_MAXG * CB_MAXD];
        gather_x(dim, ng, xyz, cell_nodes + (int64_t)ng * cell, x);
        const double lam = Ecell[cell] * lam1, mu = Ecell[cell] * mu1;
        const int rc = dim == 3 ? elast_cell(3, nb, ng, nq, dN, dM, w, x, lam, mu, Ke, ev, sv)
                                : elast_cell(2, nb, ng, nq, dN, dM, w, x, lam, mu, Ke, ev, sv);
        if (rc) {
#pragma omp atomic write
          status = -1;
          continue;
        }
        if (scatter_n(nd, cell_dofs + (int64_t)nd * cell, Ke, colptr, rowval, nzval)) {
#pragma omp atomic write
          status = -2;
        }
      }
    }
    free(Ke);
    free(ev);
    free(sv);
  }
  return status;
}

/* Compliance adjoint for the SIMP element E = emin + θ^p (1 − emin) on isoparametric Q1
 * vector cells (geometry base = function base, tables dM, w): per cell the energy
 * u_eᵀK⁰_e u_e = Σ_q (ε(u) ⊡ C⁰ ⊡ ε(u)) dV and its derivative with respect to the node
 * coordinates,
 *   dJ/dθ_e = −p θ^(p−1) (1 − emin) u_eᵀK⁰_e u_e,
 *   dJ/dx_b,m = −E_e Σ_q dV [ψ (∇M_b)_m − 2 Σ_ik (∇u)_im σ_ik (∇M_b)_k],
 * accumulated per node over the cells of one colour at a time.  dJdX[node*dim + m] is
 * zeroed first.  Returns 0, -1 on det(J) <= 0, -3 on sizes beyond the limits. */
int cb_compliance_adjoint(int dim, int ng, int nq, const double* dM, const double* w, int64_t ncells,
                          int64_t nnodes, const double* xyz, const int32_t* cell_nodes, const int32_t* cell_dofs,
                          const double* theta, double nu, double emin, double penal, const double* u,
                          int ncolors, const int64_t* color_ptr, const int32_t* color_cells, int nthreads,
                          double* dJdtheta, double* dJdX) {
  if (!check_sizes(dim, ng, ng, nq, dim)) return -3;
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  int status = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nnodes * dim; ++i) dJdX[i] = 0.0;
  for (int col = 0; col < ncolors; ++col) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t s = color_ptr[col]; s < color_ptr[col + 1]; ++s) {
      const int32_t cell = color_cells[s];
      const int32_t* nodes = cell_nodes + (int64_t)ng * cell;
      const int32_t* dofs = cell_dofs + (int64_t)ng * dim * cell;
      double x[CB_MAXG * CB_MAXD], ue[CB_MAXG][CB_MAXD], Ji[CB_MAXD][CB_MAXD], g[CB_MAXB][CB_MAXD];
      double gx[CB_MAXG][CB_MAXD] = {{0}};
      gather_x(dim, ng, xyz, nodes, x);
      for (int a = 0; a < ng; ++a)
        for (int c = 0; c < dim; ++c) ue[a][c] = u[dofs[a * dim + c]];
      double energy = 0.0;
      int bad = 0;
      for (int q = 0; q < nq; ++q) {
        const double det = reinit_q(dim, ng, ng, x, dM + (int64_t)q * ng * dim, dM + (int64_t)q * ng * dim, Ji, g);
        if (!(det > 0.0)) { bad = 1; break; }
        const double dV = det * w[q];
        double gu[CB_MAXD][CB_MAXD] = {{0}}, sig[CB_MAXD][CB_MAXD], tr = 0.0, psi = 0.0;
        for (int a = 0; a < ng; ++a)
          for (int i = 0; i < dim; ++i)
            for (int k = 0; k < dim; ++k) gu[i][k] += ue[a][i] * g[a][k];
        for (int i = 0; i < dim; ++i) tr += gu[i][i];
        for (int i = 0; i < dim; ++i)
          for (int k = 0; k < dim; ++k) {
            const double e = 0.5 * (gu[i][k] + gu[k][i]);
            sig[i][k] = 2.0 * mu * e + (i == k ? lam * tr : 0.0);
            psi += e * sig[i][k];
          }
        energy += psi * dV;
        for (int b = 0; b < ng; ++b)
          for (int m = 0; m < dim; ++m) {
            double acc = psi * g[b][m];
            for (int i = 0; i < dim; ++i)
              for (int k = 0; k < dim; ++k) acc -= 2.0 * gu[i][m] * sig[i][k] * g[b][k];
            gx[b][m] += acc * dV;
          }
      }
      if (bad) {
#pragma omp atomic write
        status = -1;
        continue;
      }
      const double th = theta[cell], E = emin + pow(th, penal) * (1.0 - emin);
      dJdtheta[cell] = -penal * pow(th, penal - 1.0) * (1.0 - emin) * energy;
      for (int b = 0; b < ng; ++b)
        for (int m = 0; m < dim; ++m) dJdX[(int64_t)nodes[b] * dim + m] -= E * gx[b][m];
    }
  }
  return status;
}

/* fᵀu (J = compliance) in a fixed sequential order. */
double cb_dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Starts the OpenMP pool with nthreads threads (a parallel sum), so that no timed call pays
 * for the threads' creation. */
double cb_warm(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < 1 << 16; ++i) s += (double)i;
  return s;
}
