// Boundary loads (FacetValues traction) and the compliance adjoint.
//
// Traction: f[dofs] += ∫_Γ t · N dΓ over a facet set, dΓ = |weighted normal| w
// (src/FEValues/FacetValues.jl:95-156, facet_integrals.jl:122-131, :191-202; the
// element routine of docs/src/literate-tutorials/topology_optimization.jl:369-381).
// One lane per facet computes its local fe into a buffer; the buffer is then summed
// into f per dof in facet-set order (a host-built dof -> contributions table), so
// the result does not depend on scheduling.
//
// Compliance adjoint (no counterpart in the reference): for K(θ, X) u = f with f
// independent of θ and X, J = fᵀu has
//   dJ/dθ_e = −uᵀ (∂K/∂θ_e) u = −p θ_e^{p−1} (1 − E_min) u_eᵀ K⁰_e u_e
//   dJ/dx_b = −Σ_e E_e Σ_q w_q detJ_q (ψ_q I − 2 ∇u_qᵀ σ_q) ∇M_b(ξ_q)
// (ψ = ε:σ, σ = C⁰:ε with E = 1; the last line is the derivative of
// detJ·ε:C:ε under x_b -> x_b + δ, with δJ = δ ⊗ ∂M_b/∂ξ).  Ke is never stored:
// both kernels recompute the element quantities per cell.  The nodal gradient is
// scattered colour by colour (no atomics), the per-cell one needs no scatter.
#include <cmath>
#include <map>

#include "fem_internal.h"

namespace fem {

struct FacetTabs {
  int nfq = 0, nft = 0, nbf = 0, ngeo = 0, rdim = 0;
  const double* N = nullptr;    // [nft][nfq][nbf]
  const double* dM = nullptr;   // [nft][nfq][ngeo][rdim]
  const double* w = nullptr;    // [nft][nfq]
};

__device__ __forceinline__ void cross3(const double* u, const double* v, double* o) {
  o[0] = u[1] * v[2] - u[2] * v[1];
  o[1] = u[2] * v[0] - u[0] * v[2];
  o[2] = u[0] * v[1] - u[1] * v[0];
}

// One lane per facet: fe[facet][I] (I node-major, component-minor).
__global__ void k_facet_fe(int64_t nf, const int32_t* __restrict__ fcell, const int32_t* __restrict__ fnum,
                           const int32_t* __restrict__ cell_nodes, int npc, const double* __restrict__ X0,
                           const double* __restrict__ X1, const double* __restrict__ X2, int sdim, int nfq, int nbf,
                           int ngeo, const double* __restrict__ tN, const double* __restrict__ tdM,
                           const double* __restrict__ tw, double t0, double t1, double t2, int vdim,
                           double* __restrict__ fe_out) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const int32_t cell = fcell[k], fc = fnum[k];
  double x[8][3];
  for (int a = 0; a < ngeo; ++a) {
    const int32_t nd = cell_nodes[(int64_t)cell * npc + a];
    x[a][0] = X0[nd];
    x[a][1] = X1[nd];
    x[a][2] = sdim == 3 ? X2[nd] : 0.0;
  }
  const int ndpc = nbf * vdim;
  double* fe = fe_out + k * ndpc;
  for (int I = 0; I < ndpc; ++I) fe[I] = 0.0;
  const double tv[3] = {t0, t1, t2};
  for (int q = 0; q < nfq; ++q) {
    const int tq = fc * nfq + q;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < ngeo; ++a)
      for (int i = 0; i < sdim; ++i)
        for (int j = 0; j < sdim; ++j) J[i][j] += x[a][i] * tdM[(tq * ngeo + a) * sdim + j];
    double nrm;
    if (sdim == 2) {
      // facet_integrals.jl:122-131
      double n0, n1;
      switch (fc) {
        case 0: n0 = J[1][0]; n1 = -J[0][0]; break;
        case 1: n0 = J[1][1]; n1 = -J[0][1]; break;
        case 2: n0 = -J[1][0]; n1 = J[0][0]; break;
        default: n0 = -J[1][1]; n1 = J[0][1]; break;
      }
      nrm = sqrt(n0 * n0 + n1 * n1);
    } else {
      // facet_integrals.jl:191-202: column pairs (2,1) (1,3) (2,3) (3,1) (3,2) (1,2)
      const int pa[6] = {1, 0, 1, 2, 2, 0}, pb[6] = {0, 2, 2, 0, 1, 1};
      const double ca[3] = {J[0][pa[fc]], J[1][pa[fc]], J[2][pa[fc]]};
      const double cb[3] = {J[0][pb[fc]], J[1][pb[fc]], J[2][pb[fc]]};
      double n[3];
      cross3(ca, cb, n);
      nrm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    }
    const double dG = nrm * tw[tq];
    for (int a = 0; a < nbf; ++a) {
      const double Na = tN[tq * nbf + a];
      for (int d = 0; d < vdim; ++d) fe[a * vdim + d] += tv[d] * Na * dG;
    }
  }
}

// f[dof] += Σ contributions in facet order (CSR per touched dof).
__global__ void k_facet_gather(int64_t nd, const int32_t* __restrict__ dof, const int32_t* __restrict__ ptr,
                               const int32_t* __restrict__ src, const double* __restrict__ fe,
                               double* __restrict__ f) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nd) return;
  double acc = f[dof[k]];
  for (int32_t p = ptr[k]; p < ptr[k + 1]; ++p) acc += fe[src[p]];
  f[dof[k]] = acc;
}

// ---------------------------------------------------------------------------
// compliance adjoint, small vector elements (Q1 quad^2 plane strain, Q1 hex^3)
// ---------------------------------------------------------------------------
template <int DIM, int NG, int NQ>
__global__ void k_adjoint(int64_t s0, int64_t s1, const double* __restrict__ X0, const double* __restrict__ X1,
                          const double* __restrict__ X2, const int32_t* __restrict__ cnodes,
                          const int32_t* __restrict__ cdofs, const int32_t* __restrict__ ccell,
                          const double* __restrict__ theta_cellorder, const double* __restrict__ tab, double lam,
                          double mu, double emin, double penal, const double* __restrict__ u,
                          double* __restrict__ dJdtheta, double* __restrict__ dJdX) {
  const int64_t s = s0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= s1) return;
  constexpr int odN = NQ * NG, odM = odN + NQ * NG * DIM, ow = odM + NQ * NG * DIM;
  const int32_t cell = ccell[s];
  double x[NG][DIM], ue[NG][DIM];
  for (int a = 0; a < NG; ++a) {
    const int32_t nd = cnodes[s * NG + a];
    x[a][0] = X0[nd];
    x[a][1] = X1[nd];
    if constexpr (DIM == 3) x[a][2] = X2[nd];
    for (int i = 0; i < DIM; ++i) ue[a][i] = u[cdofs[s * NG * DIM + a * DIM + i]];
  }
  const double th = theta_cellorder[cell];
  const double E = emin + pow(th, penal) * (1.0 - emin);
  double energy0 = 0.0;          // u_eᵀ K⁰_e u_e
  double gx[NG][DIM] = {};       // ∂(u_eᵀ K⁰_e u_e)/∂x_b
  for (int q = 0; q < NQ; ++q) {
    double J[DIM][DIM] = {};
    for (int a = 0; a < NG; ++a)
      for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j) J[i][j] += x[a][i] * tab[odM + (q * NG + a) * DIM + j];
    double Ji[DIM][DIM];
    double detJ;
    if constexpr (DIM == 2) {
      detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      Ji[0][0] = J[1][1] / detJ; Ji[0][1] = -J[0][1] / detJ;
      Ji[1][0] = -J[1][0] / detJ; Ji[1][1] = J[0][0] / detJ;
    } else {
      const double a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
      const double a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      detJ = J[0][0] * a00 + J[1][0] * a01 + J[2][0] * a02;
      const double id = 1.0 / detJ;
      Ji[0][0] = a00 * id; Ji[0][1] = a01 * id; Ji[0][2] = a02 * id;
      Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
      Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
      Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
    const double wq = tab[ow + q];
    double g[NG][DIM];
    for (int b = 0; b < NG; ++b)
      for (int j = 0; j < DIM; ++j) {
        double acc = 0.0;
        for (int i = 0; i < DIM; ++i) acc += tab[odN + (q * NG + b) * DIM + i] * Ji[i][j];
        g[b][j] = acc;
      }
    double gu[DIM][DIM] = {};  // ∇u: gu[i][k] = Σ_a u_a,i ∂_k N_a
    for (int a = 0; a < NG; ++a)
      for (int i = 0; i < DIM; ++i)
        for (int k = 0; k < DIM; ++k) gu[i][k] += ue[a][i] * g[a][k];
    double eps[DIM][DIM], sig[DIM][DIM], tr = 0.0;
    for (int i = 0; i < DIM; ++i) tr += gu[i][i];
    for (int i = 0; i < DIM; ++i)
      for (int k = 0; k < DIM; ++k) {
        eps[i][k] = 0.5 * (gu[i][k] + gu[k][i]);
        sig[i][k] = 2.0 * mu * eps[i][k] + (i == k ? lam * tr : 0.0);
      }
    double psi = 0.0;
    for (int i = 0; i < DIM; ++i)
      for (int k = 0; k < DIM; ++k) psi += eps[i][k] * sig[i][k];
    const double dV = detJ * wq;
    energy0 += psi * dV;
    // (ψ I − 2 ∇uᵀ σ) ∇M_b   (geometry ip == function ip here, so ∇M_b = g[b])
    double Tm[DIM][DIM];
    for (int m = 0; m < DIM; ++m)
      for (int k = 0; k < DIM; ++k) {
        double acc = 0.0;
        for (int i = 0; i < DIM; ++i) acc += gu[i][m] * sig[i][k];
        Tm[m][k] = (m == k ? psi : 0.0) - 2.0 * acc;
      }
    for (int b = 0; b < NG; ++b)
      for (int m = 0; m < DIM; ++m) {
        double acc = 0.0;
        for (int k = 0; k < DIM; ++k) acc += Tm[m][k] * g[b][k];
        gx[b][m] += dV * acc;
      }
  }
  dJdtheta[cell] = -penal * pow(th, penal - 1.0) * (1.0 - emin) * energy0;
  for (int b = 0; b < NG; ++b) {
    const int32_t nd = cnodes[s * NG + b];
    for (int m = 0; m < DIM; ++m) dJdX[(int64_t)nd * DIM + m] -= E * gx[b][m];
  }
}

// Partials of uᵀ K u over the CSC pattern (K symmetric: column j is row j), one value per block.
__global__ __launch_bounds__(256) void k_energy_partial(int64_t n, const int32_t* __restrict__ colptr,
                                                        const int32_t* __restrict__ rowval,
                                                        const double* __restrict__ nzval,
                                                        const double* __restrict__ u, double* __restrict__ part) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double kj = 0.0;
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p) kj += nzval[p] * u[rowval[p]];
    acc += u[j] * kj;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void k_dot_partial(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                              double* __restrict__ part) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) acc += a[i] * b[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_zero_d(int64_t n, double* __restrict__ a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

}  // namespace fem

using namespace fem;

extern "C" {

int32_t fem_facet_traction(fem_ctx* c, int64_t nfacets, const int64_t* facets, int32_t nfq, int32_t nft,
                           const double* fN, const double* fdMdxi, const double* fw, const double* traction) {
  if (!c || nfacets < 0 || (nfacets > 0 && (!facets || !fN || !fdMdxi || !fw || !traction)) || nfq < 1 ||
      nft < 1 || nft > 6 || nfq > kMaxNFQ)
    return set_error(c, FEM_ERR_ARG, "fem_facet_traction: invalid arguments");
  if (!c->have_cv || c->ndpc == 0 || c->ncells == 0) return set_error(c, FEM_ERR_STATE, "upload CellValues and dofs first");
  if (c->rdim != c->sdim || c->ndpc != c->nbf * c->vdim || c->ngeo > 8)
    return set_error(c, FEM_ERR_ARG, "facet traction needs rdim == sdim and ndofs_per_cell == nbf * vdim");
  if (nfacets == 0) return FEM_OK;
  const int nbf = c->nbf, ngeo = c->ngeo, sdim = c->sdim, vdim = c->vdim, ndpc = c->ndpc;
  std::vector<int32_t> fc(nfacets), fnum(nfacets);
  for (int64_t k = 0; k < nfacets; ++k) {
    const int64_t cell = facets[2 * k], f = facets[2 * k + 1];
    if (cell < 1 || cell > c->ncells || f < 1 || f > nft) return set_error(c, FEM_ERR_ARG, "facet index out of range");
    fc[k] = (int32_t)(cell - 1);
    fnum[k] = (int32_t)(f - 1);
  }
  // dof -> contributions in facet order
  std::map<int32_t, std::vector<int32_t>> contrib;
  for (int64_t k = 0; k < nfacets; ++k)
    for (int I = 0; I < ndpc; ++I) contrib[c->h_cell_dofs[(int64_t)fc[k] * ndpc + I]].push_back((int32_t)(k * ndpc + I));
  std::vector<int32_t> dof, ptr(1, 0), src;
  for (auto& kv : contrib) {
    dof.push_back(kv.first);
    src.insert(src.end(), kv.second.begin(), kv.second.end());
    ptr.push_back((int32_t)src.size());
  }
  // All facet contributions are summed (zero ones included), like assemble!(f, dofs, fe).
  int32_t *d_fc = nullptr, *d_fn = nullptr, *d_cn = nullptr, *d_dof = nullptr, *d_ptr = nullptr, *d_src = nullptr;
  double *d_N = nullptr, *d_dM = nullptr, *d_w = nullptr, *d_fe = nullptr;
  const size_t nN = (size_t)nft * nfq * nbf, ndM = (size_t)nft * nfq * ngeo * sdim, nw = (size_t)nft * nfq;
  int rc;
  auto cleanup = [&]() {
    dev_free(&d_fc); dev_free(&d_fn); dev_free(&d_cn); dev_free(&d_dof); dev_free(&d_ptr); dev_free(&d_src);
    dev_free(&d_N); dev_free(&d_dM); dev_free(&d_w); dev_free(&d_fe);
  };
  if ((rc = dev_alloc(c, &d_fc, nfacets)) || (rc = dev_alloc(c, &d_fn, nfacets)) ||
      (rc = dev_alloc(c, &d_cn, c->h_cell_nodes.size())) || (rc = dev_alloc(c, &d_dof, dof.size())) ||
      (rc = dev_alloc(c, &d_ptr, ptr.size())) || (rc = dev_alloc(c, &d_src, src.size())) ||
      (rc = dev_alloc(c, &d_N, nN)) || (rc = dev_alloc(c, &d_dM, ndM)) || (rc = dev_alloc(c, &d_w, nw)) ||
      (rc = dev_alloc(c, &d_fe, (size_t)nfacets * ndpc))) {
    cleanup();
    return rc;
  }
  hipStream_t st = c->stream;
  FEM_HIP(c, hipMemcpyAsync(d_fc, fc.data(), 4 * nfacets, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_fn, fnum.data(), 4 * nfacets, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_cn, c->h_cell_nodes.data(), 4 * c->h_cell_nodes.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_dof, dof.data(), 4 * dof.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_ptr, ptr.data(), 4 * ptr.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_src, src.data(), 4 * src.size(), hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_N, fN, 8 * nN, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_dM, fdMdxi, 8 * ndM, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_w, fw, 8 * nw, hipMemcpyHostToDevice, st));
  const double t0 = traction[0], t1 = sdim > 1 ? traction[1] : 0.0, t2 = sdim > 2 ? traction[2] : 0.0;
  k_facet_fe<<<(unsigned)((nfacets + 127) / 128), 128, 0, st>>>(nfacets, d_fc, d_fn, d_cn, c->npc, c->d_x[0], c->d_x[1],
                                                              c->d_x[2], sdim, nfq, nbf, ngeo, d_N, d_dM, d_w, t0, t1,
                                                              t2, vdim, d_fe);
  k_facet_gather<<<(unsigned)((dof.size() + 255) / 256), 256, 0, st>>>((int64_t)dof.size(), d_dof, d_ptr, d_src, d_fe,
                                                                      c->d_f);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return check_hip(c, e, "facet traction");
  return FEM_OK;
}

int32_t fem_adjoint_compliance(fem_ctx* c, const double* theta, double emin, double penal, double nu, double* J_out,
                               double* dJdtheta, double* dJdX) {
  if (!c || !theta) return set_error(c, FEM_ERR_ARG, "fem_adjoint_compliance: invalid arguments");
  if (!c->tables_ready || !c->have_cv) return set_error(c, FEM_ERR_STATE, "assemble and solve the system first");
  if (!(c->vdim == c->rdim && c->sdim == c->rdim && c->nbf == c->ngeo && c->ndpc == c->nbf * c->vdim &&
        ((c->rdim == 2 && c->nqp == 4) || (c->rdim == 3 && c->nqp == 8))))
    return set_error(c, FEM_ERR_ARG, "compliance adjoint supports Q1 quad^2 (2x2 Gauss) and Q1 hex^3 (2x2x2 Gauss)");
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  const int64_t nc = c->ncells, nn = c->nnodes;
  double *d_th = nullptr, *d_dth = nullptr, *d_dx = nullptr, *d_tab = nullptr;
  int rc;
  if ((rc = dev_alloc(c, &d_th, nc)) || (rc = dev_alloc(c, &d_dth, nc)) || (rc = dev_alloc(c, &d_dx, nn * c->sdim))) {
    dev_free(&d_th); dev_free(&d_dth); dev_free(&d_dx);
    return rc;
  }
  // packed tables (N, dN, dM, w) in the layout the kernel indexes
  const int DIM = c->rdim, NG = c->ngeo, NQ = c->nqp;
  std::vector<double> tab((size_t)NQ * NG + 2 * (size_t)NQ * NG * DIM + NQ);
  for (int q = 0; q < NQ; ++q) {
    for (int b = 0; b < NG; ++b) {
      tab[q * NG + b] = c->ref.N[q][b];
      for (int d = 0; d < DIM; ++d) {
        tab[NQ * NG + (q * NG + b) * DIM + d] = c->ref.dN[q][b][d];
        tab[NQ * NG + NQ * NG * DIM + (q * NG + b) * DIM + d] = c->ref.dM[q][b][d];
      }
    }
    tab[NQ * NG + 2 * NQ * NG * DIM + q] = c->ref.w[q];
  }
  if ((rc = dev_alloc(c, &d_tab, tab.size()))) {
    dev_free(&d_th); dev_free(&d_dth); dev_free(&d_dx);
    return rc;
  }
  hipStream_t st = c->stream;
  FEM_HIP(c, hipMemcpyAsync(d_th, theta, 8 * nc, hipMemcpyHostToDevice, st));
  FEM_HIP(c, hipMemcpyAsync(d_tab, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice, st));
  k_zero_d<<<1024, 256, 0, st>>>(nn * c->sdim, d_dx);
  for (int col = 0; col < c->ncolors; ++col) {
    const int64_t s0 = c->color_ptr[col], s1 = c->color_ptr[col + 1];
    if (s1 <= s0) continue;
    const unsigned nb = (unsigned)((s1 - s0 + 127) / 128);
    if (DIM == 2)
      k_adjoint<2, 4, 4><<<nb, 128, 0, st>>>(s0, s1, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes, c->d_cdofs,
                                              c->d_ccell, d_th, d_tab, lam, mu, emin, penal, c->d_u, d_dth, d_dx);
    else
      k_adjoint<3, 8, 8><<<nb, 128, 0, st>>>(s0, s1, c->d_x[0], c->d_x[1], c->d_x[2], c->d_cnodes, c->d_cdofs,
                                              c->d_ccell, d_th, d_tab, lam, mu, emin, penal, c->d_u, d_dth, d_dx);
  }
  // J = fᵀu evaluated as fᵀu + uᵀ(f − K u) = 2 fᵀu − uᵀK u: equal at the exact solution,
  // and its error is quadratic in the solver residual (rᵀK⁻¹r) instead of linear, so J
  // is accurate to rounding even when u carries the CG tolerance (fixed-order sums).
  k_dot_partial<<<1024, 256, 0, st>>>(c->ndofs, c->d_f, c->d_u, c->d_partials);
  std::vector<double> part(1024), epart(1024);
  FEM_HIP(c, hipMemcpyAsync(part.data(), c->d_partials, 8 * 1024, hipMemcpyDeviceToHost, st));
  FEM_HIP(c, hipStreamSynchronize(st));
  k_energy_partial<<<1024, 256, 0, st>>>(c->ndofs, c->d_colptr, c->d_rowval, c->d_nzval, c->d_u, c->d_partials);
  FEM_HIP(c, hipMemcpyAsync(epart.data(), c->d_partials, 8 * 1024, hipMemcpyDeviceToHost, st));
  if (dJdtheta) FEM_HIP(c, hipMemcpyAsync(dJdtheta, d_dth, 8 * nc, hipMemcpyDeviceToHost, st));
  if (dJdX) FEM_HIP(c, hipMemcpyAsync(dJdX, d_dx, 8 * nn * c->sdim, hipMemcpyDeviceToHost, st));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dev_free(&d_th); dev_free(&d_dth); dev_free(&d_dx); dev_free(&d_tab);
  if (e != hipSuccess) return check_hip(c, e, "compliance adjoint");
  double fu = 0.0, uku = 0.0;
  for (double v : part) fu += v;
  for (double v : epart) uku += v;
  if (J_out) *J_out = 2.0 * fu - uku;
  return FEM_OK;
}

}  // extern "C"
