// Fused cell-loop assembly: gather -> CellValues reinit! -> quadrature loop ->
// Ke/fe -> deterministic colored scatter into CSC nzval and f.
//
// Replaces, per colour, the body of the threaded colored loop
// (docs/src/literate-howto/threaded_assembly.jl:232-265):
//   reinit!(cellcache, i)           iterators.jl:72-88     (gather coords/dofs)
//   reinit!(cellvalues, cellcache)  CellValues.jl:122-140  (J, detJ, J⁻¹, ∇N)
//   assemble_element!               heat_equation.jl:143-164
//   assemble!(assembler, dofs,...)  assembler.jl:322-457   (scatter)
//
// Kernel shape (one launch per colour): a block owns BLOCK consecutive colour
// slots.  Phase 1 maps one lane to one cell and keeps the whole element
// computation in registers (no redundant geometry work, ~36 wave-instructions
// per cell for Q1 hex).  The upper triangle of Ke is staged in LDS.  Phase 2
// remaps lanes to (cell, entry) so that consecutive lanes write consecutive
// rows of one column of K: one wave instruction touches the 1-2 cache lines of
// 8 columns instead of 64 scattered lines.  Cells of one colour share no dof,
// so the read-modify-write of nzval/f needs no atomics and the sum order per
// entry is the colour order: results are bitwise reproducible run to run.
#include "fem_element.h"

namespace fem {

__global__ void k_fill_zero(double* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_reset_err(DevError* e) {
  e->detj_cell = INT_MAX;
  e->pattern_cell = INT_MAX;
  e->pattern_entry = -1;
}

// The constraints handed to the element kernel for apply_local!: per-dof flag, the sorted
// prescribed dofs and their inhomogeneities (0-based), apply_zero.
struct LocalBC {
  const uint8_t* isc;
  const int32_t* pres;
  const double* g;
  int64_t np;
  int zero;
};

// apply_local!(Ke, fe, dofs, ch) (ConstraintHandler.jl:1763-1820) on one cell's symmetric Ke
// (upper triangle, tri_idx) and fe, Dirichlet constraints: fe −= Ke[:, a] g_a for each
// constrained local dof a in order (not with apply_zero), m = meandiag(Ke) before any change,
// then each constrained row and column zeroed, Ke[a, a] = m, fe[a] = g_a m (0: apply_zero).
// Products and sums unfused, in the reference's order.
template <typename T, int NB>
__device__ __forceinline__ void apply_local_cell(T* ke, T* fe, const int32_t* dofs, const LocalBC& bc) {
#pragma clang fp contract(off)
  bool con[NB];
  T gv[NB];
  bool any = false;
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    const int32_t d = dofs[a];
    con[a] = bc.isc[d] != 0;
    gv[a] = T(0);
    if (con[a]) {
      int64_t lo = 0, hi = bc.np;   // d is in pres[0, np)
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (bc.pres[mid] < d) lo = mid + 1; else hi = mid;
      }
      gv[a] = (T)bc.g[lo];
      any = true;
    }
  }
  if (!any) return;
  if (!bc.zero) {
#pragma unroll
    for (int a = 0; a < NB; ++a)
      if (con[a] && gv[a] != T(0)) {
#pragma unroll
        for (int j = 0; j < NB; ++j) fe[j] = fe[j] - gv[a] * ke[tri_idx(j < a ? j : a, j < a ? a : j, NB)];
      }
  }
  T m = T(0);
#pragma unroll
  for (int i = 0; i < NB; ++i) m = m + fabs(ke[tri_idx(i, i, NB)]);
  m = m / (T)NB;
#pragma unroll
  for (int a = 0; a < NB; ++a)
    if (con[a]) {
#pragma unroll
      for (int j = 0; j < NB; ++j) ke[tri_idx(j < a ? j : a, j < a ? a : j, NB)] = T(0);
      ke[tri_idx(a, a, NB)] = m;
      fe[a] = bc.zero ? T(0) : gv[a] * m;
    }
}

// Scalar Poisson/heat element, Q1 geometry == function interpolation.  LOCAL: apply_local!
// with `bc` on every cell before it is added.
template <typename T, int DIM, int NG, int NB, int NQ, int BLOCK, bool LOCAL>
__global__ __launch_bounds__(BLOCK) void k_poisson(
    int64_t s0, int64_t s1, const double* __restrict__ X0, const double* __restrict__ X1,
    const double* __restrict__ X2, const int32_t* __restrict__ cnodes, const int32_t* __restrict__ cdofs,
    const int32_t* __restrict__ cnz, const int32_t* __restrict__ ccell, const double* __restrict__ cparam,
    const T* __restrict__ tab, T src, double* __restrict__ nzval, double* __restrict__ f,
    DevError* __restrict__ err, LocalBC bc) {
  constexpr int TRI = NB * (NB + 1) / 2;
  __shared__ T s_ke[BLOCK * TRI];
  __shared__ T s_fe[BLOCK * NB];
  __shared__ T stab[TabLayout<DIM, NG, NB, NQ>::size];
  stage_table<T, TabLayout<DIM, NG, NB, NQ>::size>(stab, tab);

  const int64_t sb = s0 + (int64_t)blockIdx.x * BLOCK;
  const int ncell = (int)min<int64_t>(BLOCK, s1 - sb);
  const int lc = threadIdx.x;

  if (lc < ncell) {
    const int64_t s = sb + lc;
    T x[NG][DIM];
#pragma unroll
    for (int a = 0; a < NG; ++a) {
      const int32_t nd = cnodes[s * NG + a];
      x[a][0] = (T)X0[nd];
      x[a][1] = (T)X1[nd];
      if constexpr (DIM == 3) x[a][2] = (T)X2[nd];
    }
    T ke[TRI];
    T fe[NB];
    const bool bad = !poisson_cell<T, DIM, NG, NB, NQ>(x, (T)cparam[s], src, stab, ke, fe);
    if (bad) atomicMin(&err->detj_cell, ccell[s]);
    if constexpr (LOCAL) apply_local_cell<T, NB>(ke, fe, cdofs + s * NB, bc);
#pragma unroll
    for (int i = 0; i < TRI; ++i) s_ke[lc * TRI + i] = ke[i];
#pragma unroll
    for (int i = 0; i < NB; ++i) s_fe[lc * NB + i] = fe[i];
  }
  __syncthreads();

  // Phase 2: lanes = (cell, column b, row a); entry t = cell*NB*NB + b*NB + a.  Each thread
  // first loads CH table entries, then the CH nzval values they address (all independent),
  // then writes the CH sums back, so CH read-modify-writes are in flight per lane.
  constexpr int E = NB * NB;
  constexpr int CH = 16;
  const int32_t* __restrict__ nzt = cnz + sb * E;
  const int tot = ncell * E;
  for (int base = 0; base < tot; base += BLOCK * CH) {
    int32_t pp[CH];
    double old[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int t = base + k * BLOCK + threadIdx.x;
      pp[k] = t < tot ? nzt[t] : -1;
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) old[k] = pp[k] >= 0 ? nzval[pp[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int t = base + k * BLOCK + threadIdx.x;
      if (t < tot) {
        const int lcell = t / E;
        const int e = t - lcell * E;
        const int b = e / NB, a = e - b * NB;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const T v = s_ke[lcell * TRI + tri_idx(lo, hi, NB)];
        if (pp[k] >= 0) {
          if (v != T(0)) nzval[pp[k]] = old[k] + (double)v;
        } else if (v != T(0)) {
          atomicMin(&err->pattern_cell, ccell[sb + lcell]);
        }
      }
    }
  }
  constexpr int CF = (BLOCK * NB + BLOCK - 1) / BLOCK;
  const int32_t* __restrict__ dt = cdofs + sb * NB;
  int32_t dd[CF];
  double fo[CF];
#pragma unroll
  for (int k = 0; k < CF; ++k) {
    const int t = k * BLOCK + threadIdx.x;
    dd[k] = t < ncell * NB ? dt[t] : -1;
  }
#pragma unroll
  for (int k = 0; k < CF; ++k) fo[k] = dd[k] >= 0 ? f[dd[k]] : 0.0;
#pragma unroll
  for (int k = 0; k < CF; ++k) {
    const int t = k * BLOCK + threadIdx.x;
    if (dd[k] >= 0) f[dd[k]] = fo[k] + (double)s_fe[t];
  }
}

template <typename T, int DIM, int NG, int NB, int NQ>
static void pack_tables(const RefTables& r, std::vector<T>& out) {
  using L = TabLayout<DIM, NG, NB, NQ>;
  out.assign(L::size, T(0));
  for (int q = 0; q < NQ; ++q) {
    out[L::ow + q] = (T)r.w[q];
    for (int b = 0; b < NB; ++b) {
      out[L::oN + q * NB + b] = (T)r.N[q][b];
      for (int d = 0; d < DIM; ++d) out[L::odN + (q * NB + b) * DIM + d] = (T)r.dN[q][b][d];
    }
    for (int a = 0; a < NG; ++a)
      for (int d = 0; d < DIM; ++d) out[L::odM + (q * NG + a) * DIM + d] = (T)r.dM[q][a][d];
  }
}

// Trilinear hex geometry: Ferrite's node order (reference coordinates below) with
// dM_a(ξ) = s_a ⊙ [(1 + s_η η)(1 + s_ζ ζ), (1 + s_ξ ξ)(1 + s_ζ ζ), (1 + s_ξ ξ)(1 + s_η η)] / 8.
// Recovers each point's (ξ, η, ζ) from the uploaded dM table and accepts the point only if
// the whole table is reproduced to 1e-14.  qc[q] = ξ, η, ζ, ηζ, ξζ, ξη.
static bool detect_trilinear(const fem_ctx* c, std::vector<double>& qc) {
  if (c->rdim != 3 || c->ngeo != 8) return false;
  static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  qc.assign(6 * c->nqp, 0.0);
  for (int q = 0; q < c->nqp; ++q) {
    const auto& dM = c->ref.dM[q];
    // ∂M/∂ξ of nodes 0 (η = ζ = −1 signs) and 3 (η = +1): η = (d3 − d0)/(d3 + d0), etc.
    const double e0 = dM[0][0], e3 = dM[3][0], e4 = dM[4][0];
    const double f0 = dM[0][1], f1 = dM[1][1];
    if (e0 + e3 == 0.0 || e0 + e4 == 0.0 || f0 + f1 == 0.0) return false;
    const double eta = (e3 - e0) / (e3 + e0);
    const double zeta = (e4 - e0) / (e4 + e0);
    const double xi = (f1 - f0) / (f1 + f0);
    const double p[3] = {xi, eta, zeta};
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) {
        double v = sg[a][d] / 8.0;
        for (int e = 0; e < 3; ++e)
          if (e != d) v *= 1.0 + sg[a][e] * p[e];
        if (std::fabs(v - dM[a][d]) > 1e-14) return false;
      }
    double* o = &qc[6 * q];
    o[0] = xi; o[1] = eta; o[2] = zeta; o[3] = eta * zeta; o[4] = xi * zeta; o[5] = xi * eta;
  }
  return true;
}

template <int DIM, int NG, int NB, int NQ>
static int upload_tables(fem_ctx* c) {
  if (!c->tab_dirty) return FEM_OK;
  std::vector<double> td;
  std::vector<float> tf;
  pack_tables<double, DIM, NG, NB, NQ>(c->ref, td);
  pack_tables<float, DIM, NG, NB, NQ>(c->ref, tf);
  int rc;
  if ((rc = dev_alloc(c, &c->d_tab, td.size())) || (rc = dev_alloc(c, &c->d_tabf, tf.size()))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_tab, td.data(), td.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_tabf, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  std::vector<double> qc;
  c->trilinear = detect_trilinear(c, qc);
  if (c->trilinear) {
    std::vector<float> qf(qc.begin(), qc.end());
    if ((rc = dev_alloc(c, &c->d_qc, qc.size())) || (rc = dev_alloc(c, &c->d_qcf, qf.size()))) return rc;
    FEM_HIP(c, hipMemcpyAsync(c->d_qc, qc.data(), qc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->d_qcf, qf.data(), qf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    // reference integrals of the affine fast path (same quadrature as the element loop),
    // strict upper triangle in row-major order (the stored element planes), then Σ_q w_q N_a
    const int nb = c->nbf, nq = c->nqp;
    static const int mi[6] = {0, 1, 2, 0, 0, 1}, mj[6] = {0, 1, 2, 1, 2, 2};
    std::vector<double> aff;
    for (int a = 0; a < nb; ++a)
      for (int b = a + 1; b < nb; ++b)
        for (int m = 0; m < 6; ++m) {
          double v = 0.0;
          for (int q = 0; q < nq; ++q) {
            const auto& g = c->ref.dN[q];
            double t = g[a][mi[m]] * g[b][mj[m]];
            if (mi[m] != mj[m]) t += g[a][mj[m]] * g[b][mi[m]];
            v += c->ref.w[q] * t;
          }
          aff.push_back(v);
        }
    for (int a = 0; a < nb; ++a) {
      double v = 0.0;
      for (int q = 0; q < nq; ++q) v += c->ref.w[q] * c->ref.N[q][a];
      aff.push_back(v);
    }
    std::vector<float> afff(aff.begin(), aff.end());
    if ((rc = dev_alloc(c, &c->d_aff, aff.size())) || (rc = dev_alloc(c, &c->d_afff, afff.size()))) return rc;
    FEM_HIP(c, hipMemcpyAsync(c->d_aff, aff.data(), aff.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->d_afff, afff.data(), afff.size() * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
  }
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->tab_dirty = false;
  return FEM_OK;
}

template <typename T, int DIM, int NG, int NB, int NQ, int BLOCK>
static int launch_poisson(fem_ctx* c, double src, int local) {
  int rc = upload_tables<DIM, NG, NB, NQ>(c);
  if (rc) return rc;
  const T* tab = nullptr;
  if constexpr (sizeof(T) == 8) tab = c->d_tab; else tab = c->d_tabf;
  FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
  for (int col = 0; col < c->ncolors; ++col) {
    const int64_t s0 = c->color_ptr[col], s1 = c->color_ptr[col + 1];
    if (s1 <= s0) continue;
    const int64_t nb = (s1 - s0 + BLOCK - 1) / BLOCK;
    const LocalBC bc{c->d_isc, c->d_pres, c->d_g, c->np, local == 2 ? 1 : 0};
    if (local)
      k_poisson<T, DIM, NG, NB, NQ, BLOCK, true><<<dim3((unsigned)nb), dim3(BLOCK), 0, c->stream>>>(
          s0, s1, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes, c->d_cdofs, c->d_cnz, c->d_ccell,
          c->d_cparam, tab, (T)src, c->d_nzval, c->d_f, c->d_err, bc);
    else
      k_poisson<T, DIM, NG, NB, NQ, BLOCK, false><<<dim3((unsigned)nb), dim3(BLOCK), 0, c->stream>>>(
          s0, s1, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes, c->d_cdofs, c->d_cnz, c->d_ccell,
          c->d_cparam, tab, (T)src, c->d_nzval, c->d_f, c->d_err, bc);
    c->timings.assemble_launches++;
  }
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->timings.assemble_kernel_ms = event_ms(c);
  return FEM_OK;
}

// After a launch sequence: translate the device error word into Ferrite's errors.
static int check_assembly_errors(fem_ctx* c) {
  DevError e{};
  FEM_HIP(c, hipMemcpyAsync(&e, c->d_err, sizeof(DevError), hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  if (e.detj_cell != INT_MAX) {
    c->err_cell = (int64_t)e.detj_cell + 1;
    char buf[160];
    snprintf(buf, sizeof buf, "det(J) is not positive in cell %lld", (long long)c->err_cell);
    return set_error(c, FEM_ERR_DETJ, buf);
  }
  if (e.pattern_cell != INT_MAX) {
    c->err_cell = (int64_t)e.pattern_cell + 1;
    // first missing entry of that cell, found on the host from the cell's dofs
    const int nb = c->ndpc;
    const int32_t* cd = &c->h_cell_dofs[(int64_t)e.pattern_cell * nb];
    std::vector<int32_t> cp(c->n + 1);
    FEM_HIP(c, hipMemcpy(cp.data(), c->d_colptr, sizeof(int32_t) * (c->n + 1), hipMemcpyDeviceToHost));
    for (int b = 0; b < nb && c->err_row < 0; ++b) {
      const int32_t col = cd[b];
      std::vector<int32_t> rows(cp[col + 1] - cp[col]);
      FEM_HIP(c, hipMemcpy(rows.data(), c->d_rowval + cp[col], sizeof(int32_t) * rows.size(), hipMemcpyDeviceToHost));
      for (int a = 0; a < nb; ++a) {
        bool found = false;
        for (int32_t r : rows) found |= (r == cd[a]);
        if (!found) { c->err_row = cd[a] + 1; c->err_col = col + 1; break; }
      }
    }
    char buf[512];
    snprintf(buf, sizeof buf,
             "You are trying to assemble values in to K[%lld, %lld], but K[%lld, %lld] is missing in the "
             "sparsity pattern. Make sure you have called `K = allocate_matrix(dh)` or `K = "
             "allocate_matrix(dh, ch)` if you have affine constraints.",
             (long long)c->err_row, (long long)c->err_col, (long long)c->err_row, (long long)c->err_col);
    return set_error(c, FEM_ERR_PATTERN, buf);
  }
  return FEM_OK;
}

int assemble_elasticity(fem_ctx* c, const fem_assemble_cfg* cfg);

}  // namespace fem

using namespace fem;

extern "C" int32_t fem_assemble(fem_ctx* c, const fem_assemble_cfg* cfg) {
  if (!c || !cfg) return FEM_ERR_ARG;
  if (!c->have_cv) return set_error(c, FEM_ERR_STATE, "upload CellValues tables first");
  if (!c->have_param) return set_error(c, FEM_ERR_STATE, "upload the per-cell parameter first");
  FEM_HIP(c, hipSetDevice(c->device));
  c->err_cell = c->err_row = c->err_col = -1;
  c->timings.assemble_launches = 0;
  c->timings.element_ms = c->timings.scatter_ms = 0.0;
  k_reset_err<<<1, 1, 0, c->stream>>>(c->d_err);
  const bool mixed = cfg->precision == FEM_MIXED;
  const int local = cfg->local_constraints;
  if (local) {   // apply_local! on every cell: the colored scatter with the uploaded constraints
    if (!(local == 1 || local == 2)) return set_error(c, FEM_ERR_ARG, "local_constraints must be 0, 1 or 2");
    if (cfg->kernel != FEM_KERNEL_POISSON)
      return set_error(c, FEM_ERR_ARG, "apply_local! during assembly is provided for the Poisson kernel");
    if (!c->d_isc) return set_error(c, FEM_ERR_STATE, "upload constraints first (fem_constraints_upload)");
    if (c->naff > 0)
      return set_error(c, FEM_ERR_ARG, "can not condense constraint without the global matrix and vector");
  }
  // Scalar Q1 Poisson: owner-computes gather (each nzval/f entry written once, no fill
  // needed) unless the colored scatter is asked for.
  if (cfg->kernel == FEM_KERNEL_POISSON && c->tune.asm_mode != 1 && !local) {
    if (c->vdim != 1 || c->ndpc != c->nbf) return set_error(c, FEM_ERR_ARG, "Poisson kernel needs a scalar field");
    if (gather_supported(c)) {
      int rg = c->rdim == 3 ? upload_tables<3, 8, 8, 8>(c) : upload_tables<2, 4, 4, 4>(c);
      if (rg) return rg;
      c->timings.fill_ms = 0.0;
      rg = gather_poisson(c, mixed, cfg->source, cfg->fillzero != 0, c->d_tab, c->d_tabf);
      if (rg) return rg;
      ++c->nzval_version;
      elem_op_note_assembly(c, 1, !mixed, cfg->fillzero != 0);
      return check_assembly_errors(c);
    }
    if (c->tune.asm_mode == 2)
      return set_error(c, FEM_ERR_ARG, "gather assembly needs a scalar Q1 quad/hex with 2x2(x2) Gauss");
  }
  if (cfg->kernel == FEM_KERNEL_ELASTICITY) {
    // vector Q2 hex: owner-computes (element matrices, then a gather per dof node)
    bool done = false;
    int rq = assemble_elasticity_gather(c, cfg, &done);
    if (rq) return rq;
    if (done) {
      ++c->nzval_version;
      elem_op_note_assembly(c, 3, !mixed, cfg->fillzero != 0);
      return check_assembly_errors(c);
    }
  }
  int rc = build_tables(c);
  if (rc) return rc;
  if (cfg->fillzero) {
    FEM_HIP(c, hipEventRecord(c->ev0, c->stream));
    k_fill_zero<<<2048, 256, 0, c->stream>>>(c->d_nzval, c->nnz);
    k_fill_zero<<<512, 256, 0, c->stream>>>(c->d_f, c->ndofs);
    FEM_HIP(c, hipGetLastError());
    FEM_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->timings.fill_ms = event_ms(c);
  }
  if (cfg->kernel == FEM_KERNEL_POISSON) {
    if (c->vdim != 1 || c->ndpc != c->nbf) return set_error(c, FEM_ERR_ARG, "Poisson kernel needs a scalar field");
    const int key = c->rdim * 1000 + c->ngeo * 100 + c->nbf * 10 + (c->nqp == 8 ? 8 : c->nqp);
    const int bs = c->tune.asm_block;
    if (c->rdim == 3 && c->ngeo == 8 && c->nbf == 8 && c->nqp == 8 && c->npc == 8 && c->sdim == 3) {
      if (mixed) {
        rc = bs == 64 ? launch_poisson<float, 3, 8, 8, 8, 64>(c, cfg->source, local)
           : bs == 256 ? launch_poisson<float, 3, 8, 8, 8, 256>(c, cfg->source, local)
                       : launch_poisson<float, 3, 8, 8, 8, 128>(c, cfg->source, local);
      } else {
        rc = bs == 128 ? launch_poisson<double, 3, 8, 8, 8, 128>(c, cfg->source, local)
           : bs == 256 ? launch_poisson<double, 3, 8, 8, 8, 256>(c, cfg->source, local)
                       : launch_poisson<double, 3, 8, 8, 8, 64>(c, cfg->source, local);
      }
    } else if (c->rdim == 2 && c->ngeo == 4 && c->nbf == 4 && c->nqp == 4 && c->npc == 4 && c->sdim == 2) {
      rc = mixed ? launch_poisson<float, 2, 4, 4, 4, 128>(c, cfg->source, local)
                 : launch_poisson<double, 2, 4, 4, 4, 128>(c, cfg->source, local);
    } else {
      (void)key;
      return set_error(c, FEM_ERR_ARG, "no Poisson kernel for this element/quadrature combination "
                                       "(supported: Q1 quad 2x2 Gauss, Q1 hex 2x2x2 Gauss)");
    }
  } else if (cfg->kernel == FEM_KERNEL_ELASTICITY) {
    rc = assemble_elasticity(c, cfg);
  } else {
    return set_error(c, FEM_ERR_ARG, "unknown kernel id");
  }
  if (rc) return rc;
  ++c->nzval_version;
  // the Q1 quad^2 elasticity kernel stores its element matrices as planes as well
  const bool q1v = cfg->kernel == FEM_KERNEL_ELASTICITY && c->rdim == 2 && c->ndpc == 8 && c->ep_cells == c->ncells;
  elem_op_note_assembly(c, q1v ? 2 : 0, !mixed, cfg->fillzero != 0);
  return check_assembly_errors(c);
}
