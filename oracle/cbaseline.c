/*
 * Oracle (TEST INFRASTRUCTURE ONLY): C restatement of Ferrite.jl's threaded
 * colored assembly loop for Q1 Poisson, plus apply! and a Jacobi-PCG solve, used
 * as the timed CPU baseline ("Ferrite-algorithm CPU restatement") and checked
 * against the Python oracle by the CPU tests.  Never linked into the product.
 *
 * Follows, per cell (docs/src/literate-howto/threaded_assembly.jl:232-265, with
 * OpenMP tasks over the cells of each colour, colours in sequence):
 *   reinit!(cellcache)            src/iterators.jl:72-88      gather coords, dofs
 *   reinit!(cellvalues)           src/FEValues/CellValues.jl:122-140
 *     J = Σ x_j ⊗ dMdξ_j          GeometryMapping.jl:124-130
 *     detJ > 0, detJdV = detJ w   CellValues.jl:110-116
 *     dNdx = dNdξ ⋅ J⁻¹           FunctionValues.jl:163, :186-192
 *   Ke += k (∇Nᵢ·∇Nⱼ) dΩ, fe += s Nᵢ dΩ   heat_equation.jl:143-164
 *   assemble!(A, dofs, Ke, fe)    src/assembler.jl:322-457: f[dofs] += fe; sort the
 *                                 dofs with a permutation; per sorted column merge-walk
 *                                 rowval against the sorted rows (binary search when
 *                                 the column is long), skipping zero values
 * then apply!(K, f, ch)           src/Dofs/ConstraintHandler.jl:710-747 (g = 0)
 * and u = K \ f replaced by Jacobi-preconditioned CG (the GPU path's solver).
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
