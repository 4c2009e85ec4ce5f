// Context, uploads of Ferrite's host-side setup (grid, dofs, pattern, coloring,
// CellValues tables, constraints) into a device-resident layout, and downloads.
//
// Device layout (HBM):
//   coordinates       SoA x[nnodes], y[nnodes], z[nnodes] (FP64)
//   cells             permuted into color-major order once: slot s of color c
//                     holds cell color_cells[s]; per slot the tables
//                     cnodes[s][npc], cdofs[s][ndpc] and cnz[s][ndpc*ndpc]
//                     (the nzval position of K[dofs[a], dofs[b]] at entry b*ndpc+a)
//                     are contiguous, so one colour's launch streams them.
//   K                 CSC colptr/rowval (int32, 0-based), nzval (FP64)
#include <algorithm>
#include <cstring>

#include "fem_internal.h"

namespace fem {

int set_error(fem_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int check_hip(fem_ctx* c, hipError_t e, const char* what) {
  return set_error(c, FEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

float event_ms(fem_ctx* c) {
  float ms = 0.f;
  (void)hipEventSynchronize(c->ev1);
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  return ms;
}

// nz position of (row, col) by binary search in the column; -1 if absent.
__device__ __forceinline__ int32_t find_nz(const int32_t* __restrict__ colptr,
                                           const int32_t* __restrict__ rowval, int32_t row,
                                           int32_t col) {
  int32_t lo = colptr[col], hi = colptr[col + 1];
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t r = rowval[mid];
    if (r < row) lo = mid + 1; else hi = mid;
  }
  return (lo < colptr[col + 1] && rowval[lo] == row) ? lo : -1;
}

__global__ void k_build_cnz(int64_t nslots, int nb, const int32_t* __restrict__ cdofs,
                            const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowval,
                            int32_t* __restrict__ cnz) {
  const int64_t total = nslots * nb * nb;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / (nb * nb);
    const int e = (int)(t - s * nb * nb);
    const int b = e / nb, a = e - b * nb;
    const int32_t col = cdofs[s * nb + b];
    const int32_t row = cdofs[s * nb + a];
    cnz[t] = find_nz(colptr, rowval, row, col);
  }
}

__global__ void k_diag_pos(int64_t n, const int32_t* __restrict__ colptr,
                           const int32_t* __restrict__ rowval, int32_t* __restrict__ diag) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    diag[j] = find_nz(colptr, rowval, (int32_t)j, (int32_t)j);
}

__global__ void k_gather_param(int64_t nslots, const int32_t* __restrict__ ccell,
                               const double* __restrict__ param, double* __restrict__ out) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nslots;
       s += (int64_t)gridDim.x * blockDim.x)
    out[s] = param[ccell[s]];
}

static int grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 65535 * 16);
}

// Permuted per-slot tables once mesh, dofs, pattern and colors are all present.
int build_tables(fem_ctx* c) {
  if (c->tables_ready) return FEM_OK;
  if (c->ncells == 0 || c->ndpc == 0 || c->n == 0 || c->ncolors == 0)
    return set_error(c, FEM_ERR_STATE, "assemble needs mesh, dofs, pattern and colors uploaded first");
  if (c->n != c->ndofs) return set_error(c, FEM_ERR_STATE, "pattern size does not match ndofs");
  const int64_t ns = c->ncells;
  std::vector<int32_t> h(ns * (int64_t)std::max(c->npc, c->ndpc));
  // nodes
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t cell = c->h_color_cells[s];
    std::memcpy(&h[s * c->npc], &c->h_cell_nodes[(int64_t)cell * c->npc], sizeof(int32_t) * c->npc);
  }
  int rc;
  if ((rc = dev_alloc(c, &c->d_cnodes, ns * c->npc))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_cnodes, h.data(), sizeof(int32_t) * ns * c->npc, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t cell = c->h_color_cells[s];
    std::memcpy(&h[s * c->ndpc], &c->h_cell_dofs[(int64_t)cell * c->ndpc], sizeof(int32_t) * c->ndpc);
  }
  if ((rc = dev_alloc(c, &c->d_cdofs, ns * c->ndpc))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_cdofs, h.data(), sizeof(int32_t) * ns * c->ndpc, hipMemcpyHostToDevice, c->stream));
  if ((rc = dev_alloc(c, &c->d_cnz, ns * (int64_t)c->ndpc * c->ndpc))) return rc;
  const int64_t total = ns * (int64_t)c->ndpc * c->ndpc;
  k_build_cnz<<<grid_for(total), 256, 0, c->stream>>>(ns, c->ndpc, c->d_cdofs, c->d_colptr, c->d_rowval, c->d_cnz);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->tables_ready = true;
  return FEM_OK;
}

// The context's pattern bookkeeping once colptr / rowval / nzval / diag are on the device
// (fem_pattern_upload, fem_pattern_build): diagonal positions (Ferrite's patterns always store
// the diagonal; apply! writes K[d,d]), sizes, and the tables built on the old pattern dropped.
int pattern_finish(fem_ctx* c, int64_t n, int64_t nnz, const int32_t* hcolptr) {
  k_diag_pos<<<grid_for(n), 256, 0, c->stream>>>(n, c->d_colptr, c->d_rowval, c->d_diag);
  FEM_HIP(c, hipGetLastError());
  std::vector<int32_t> dg(n);
  FEM_HIP(c, hipMemcpyAsync(dg.data(), c->d_diag, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t j = 0; j < n; ++j)
    if (dg[j] < 0) return set_error(c, FEM_ERR_ARG, "pattern lacks a diagonal entry");
  c->n = n; c->nnz = nnz;
  c->pattern_sym = -1;
  c->max_row_nnz = 0;
  for (int64_t j = 0; j < n; ++j) c->max_row_nnz = std::max<int>(c->max_row_nnz, hcolptr[j + 1] - hcolptr[j]);
  free_solver(c);
  free_q2_tables(c);
  c->ea_kind = 0;
  c->tables_ready = false;
  c->gather_ready = false;
  c->affine_mesh = -1;
  return FEM_OK;
}

}  // namespace fem

using namespace fem;

extern "C" {

int32_t fem_ctx_create(int32_t device, fem_ctx** out) {
  if (!out) return FEM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FEM_ERR_HIP;
  if (device < 0 || device >= ndev) return FEM_ERR_ARG;
  fem_ctx* c = new fem_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FEM_ERR_HIP;
  }
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  (void)hipEventCreate(&c->evm);
  if (dev_alloc(c, &c->d_err, 1) || dev_alloc(c, &c->d_scalar, 16) ||
      dev_alloc(c, &c->d_partials, kReduceBlocks * 4)) {
    fem_ctx_destroy(c);
    return FEM_ERR_OOM;
  }
  c->npartials = kReduceBlocks * 4;
  // the allocation-free apply!'s scratch address on this device (once per device); a failure
  // only disables that entry point, which then reports it
  (void)noalloc_scratch_init(device);
  *out = c;
  return FEM_OK;
}

void fem_ctx_destroy(fem_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_solver(c);
  for (auto& p : c->d_x) dev_free(&p);
  dev_free(&c->d_colptr); dev_free(&c->d_rowval); dev_free(&c->d_diag);
  dev_free(&c->d_nzval);
  dev_free(&c->d_ccell); dev_free(&c->d_cnodes); dev_free(&c->d_cdofs); dev_free(&c->d_cnz);
  dev_free(&c->d_cparam); dev_free(&c->d_param_cell); dev_free(&c->d_f); dev_free(&c->d_u);
  dev_free(&c->d_gcnodes); dev_free(&c->d_d2c_ptr); dev_free(&c->d_d2c); dev_free(&c->d_gslot);
  dev_free(&c->d_keb64); dev_free(&c->d_feb64); dev_free(&c->d_keb32); dev_free(&c->d_feb32);
  free_q2_tables(c);
  dev_free(&c->d_gcdofs); dev_free(&c->d_ep64); dev_free(&c->d_ep32);
  dev_free(&c->d_pres); dev_free(&c->d_g); dev_free(&c->d_isc);
  dev_free(&c->d_affidx); dev_free(&c->d_afd); dev_free(&c->d_acp); dev_free(&c->d_acd); dev_free(&c->d_acv);
  dev_free(&c->d_acb); dev_free(&c->d_slp); dev_free(&c->d_sld); dev_free(&c->d_slv);
  dev_free(&c->d_scalar); dev_free(&c->d_partials); dev_free(&c->d_err);
  dev_free(&c->d_tab); dev_free(&c->d_tabf); dev_free(&c->d_qc); dev_free(&c->d_qcf);
  dev_free(&c->d_aff); dev_free(&c->d_afff);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evm) (void)hipEventDestroy(c->evm);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* fem_last_error(const fem_ctx* c) { return c ? c->err.c_str() : "null context"; }

int32_t fem_last_error_info(const fem_ctx* c, int64_t* cell, int64_t* row, int64_t* col) {
  if (!c) return FEM_ERR_ARG;
  if (cell) *cell = c->err_cell;
  if (row) *row = c->err_row;
  if (col) *col = c->err_col;
  return FEM_OK;
}

int32_t fem_mesh_upload(fem_ctx* c, int32_t sdim, int64_t nnodes, const double* xyz, int64_t ncells,
                        int32_t npc, const int64_t* cell_nodes) {
  if (!c || !xyz || !cell_nodes || sdim < 1 || sdim > 3 || nnodes <= 0 || ncells <= 0 || npc <= 0 || npc > kMaxNG)
    return set_error(c, FEM_ERR_ARG, "fem_mesh_upload: invalid arguments");
  if (nnodes >= INT32_MAX || ncells >= INT32_MAX) return set_error(c, FEM_ERR_ARG, "mesh too large for int32 tables");
  FEM_HIP(c, hipSetDevice(c->device));
  c->h_cell_nodes.resize(ncells * (int64_t)npc);
  for (int64_t i = 0; i < ncells * (int64_t)npc; ++i) {
    const int64_t v = cell_nodes[i];
    if (v < 1 || v > nnodes) return set_error(c, FEM_ERR_ARG, "cell node id out of range");
    c->h_cell_nodes[i] = (int32_t)(v - 1);
  }
  std::vector<double> soa(nnodes);
  for (int d = 0; d < 3; ++d) dev_free(&c->d_x[d]);
  for (int d = 0; d < sdim; ++d) {
    for (int64_t i = 0; i < nnodes; ++i) soa[i] = xyz[i * sdim + d];
    int rc = dev_alloc(c, &c->d_x[d], nnodes);
    if (rc) return rc;
    FEM_HIP(c, hipMemcpyAsync(c->d_x[d], soa.data(), sizeof(double) * nnodes, hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
  }
  c->sdim = sdim; c->nnodes = nnodes; c->ncells = ncells; c->npc = npc;
  free_solver(c);   // the solver's operator copies cache geometry (cell corners, lattice map)
  free_q2_tables(c);
  dev_free(&c->d_d2c_ptr); dev_free(&c->d_d2c); dev_free(&c->d_gcdofs);
  c->ea_kind = 0;
  c->tables_ready = false;
  c->gather_ready = false;
  c->affine_mesh = -1;
  return FEM_OK;
}

int32_t fem_dofs_upload(fem_ctx* c, int64_t ndofs, int32_t ndpc, const int64_t* cell_dofs) {
  if (!c || !cell_dofs || ndofs <= 0 || ndpc <= 0 || ndpc > 3 * kMaxNB || c->ncells == 0)
    return set_error(c, FEM_ERR_ARG, "fem_dofs_upload: invalid arguments (upload the mesh first)");
  if (ndofs >= INT32_MAX) return set_error(c, FEM_ERR_ARG, "too many dofs for int32 tables");
  const int64_t m = c->ncells * (int64_t)ndpc;
  c->h_cell_dofs.resize(m);
  for (int64_t i = 0; i < m; ++i) {
    const int64_t v = cell_dofs[i];
    if (v < 1 || v > ndofs) return set_error(c, FEM_ERR_ARG, "cell dof out of range");
    c->h_cell_dofs[i] = (int32_t)(v - 1);
  }
  c->ndofs = ndofs; c->ndpc = ndpc;
  int rc;
  if ((rc = dev_alloc(c, &c->d_f, ndofs))) return rc;
  if ((rc = dev_alloc(c, &c->d_u, ndofs))) return rc;
  FEM_HIP(c, hipMemsetAsync(c->d_f, 0, sizeof(double) * ndofs, c->stream));
  FEM_HIP(c, hipMemsetAsync(c->d_u, 0, sizeof(double) * ndofs, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  free_solver(c);   // operator copies are built on the dof numbering
  free_q2_tables(c);
  dev_free(&c->d_d2c_ptr); dev_free(&c->d_d2c); dev_free(&c->d_gcdofs);
  c->ea_kind = 0;
  c->tables_ready = false;
  c->gather_ready = false;
  c->affine_mesh = -1;
  return FEM_OK;
}

int32_t fem_pattern_upload(fem_ctx* c, int64_t n, const int64_t* colptr, const int64_t* rowval) {
  if (!c || !colptr || !rowval || n <= 0) return set_error(c, FEM_ERR_ARG, "fem_pattern_upload: invalid arguments");
  const int64_t nnz = colptr[n] - 1;
  if (colptr[0] != 1 || nnz < 0 || nnz >= INT32_MAX) return set_error(c, FEM_ERR_ARG, "invalid colptr");
  if (c->ndofs && n != c->ndofs) return set_error(c, FEM_ERR_ARG, "pattern size does not match ndofs");
  std::vector<int32_t> cp(n + 1), rv(nnz);
  for (int64_t j = 0; j <= n; ++j) cp[j] = (int32_t)(colptr[j] - 1);
  for (int64_t j = 0; j < n; ++j)
    if (cp[j + 1] < cp[j]) return set_error(c, FEM_ERR_ARG, "colptr not monotone");
  for (int64_t p = 0; p < nnz; ++p) {
    const int64_t r = rowval[p];
    if (r < 1 || r > n) return set_error(c, FEM_ERR_ARG, "rowval out of range");
    rv[p] = (int32_t)(r - 1);
  }
  int rc;
  if ((rc = dev_alloc(c, &c->d_colptr, n + 1)) || (rc = dev_alloc(c, &c->d_rowval, nnz)) ||
      (rc = dev_alloc(c, &c->d_nzval, nnz)) || (rc = dev_alloc(c, &c->d_diag, n)))
    return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_colptr, cp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(c->d_rowval, rv.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemsetAsync(c->d_nzval, 0, sizeof(double) * nnz, c->stream));
  return pattern_finish(c, n, nnz, cp.data());
}

int32_t fem_colors_upload(fem_ctx* c, int32_t ncolors, const int64_t* color_ptr, const int64_t* color_cells) {
  if (!c || ncolors <= 0 || !color_ptr || !color_cells || c->ncells == 0)
    return set_error(c, FEM_ERR_ARG, "fem_colors_upload: invalid arguments (upload the mesh first)");
  if (color_ptr[0] != 0 || color_ptr[ncolors] != c->ncells)
    return set_error(c, FEM_ERR_ARG, "colors must cover every cell exactly once");
  std::vector<uint8_t> seen(c->ncells, 0);
  c->h_color_cells.resize(c->ncells);
  for (int64_t s = 0; s < c->ncells; ++s) {
    const int64_t cell = color_cells[s];
    if (cell < 1 || cell > c->ncells || seen[cell - 1]) return set_error(c, FEM_ERR_ARG, "colors must cover every cell exactly once");
    seen[cell - 1] = 1;
    c->h_color_cells[s] = (int32_t)(cell - 1);
  }
  c->color_ptr.assign(color_ptr, color_ptr + ncolors + 1);
  for (int i = 0; i < ncolors; ++i)
    if (c->color_ptr[i + 1] < c->color_ptr[i]) return set_error(c, FEM_ERR_ARG, "color_ptr not monotone");
  int rc = dev_alloc(c, &c->d_ccell, c->ncells);
  if (rc) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_ccell, c->h_color_cells.data(), sizeof(int32_t) * c->ncells, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->ncolors = ncolors;
  c->tables_ready = false;
  c->have_param = false;
  return FEM_OK;
}

int32_t fem_cellvalues_upload(fem_ctx* c, int32_t rdim, int32_t nqp, int32_t nbf, int32_t ngeo, int32_t vdim,
                              const double* N, const double* dNdxi, const double* dMdxi, const double* w) {
  if (!c || !N || !dNdxi || !dMdxi || !w || rdim < 1 || rdim > 3 || nqp < 1 || nqp > kMaxQP || nbf < 1 ||
      nbf > kMaxNB || ngeo < 1 || ngeo > kMaxNG || vdim < 1 || vdim > 3)
    return set_error(c, FEM_ERR_ARG, "fem_cellvalues_upload: invalid arguments");
  c->ref = RefTables{};
  for (int q = 0; q < nqp; ++q) {
    c->ref.w[q] = w[q];
    for (int i = 0; i < nbf; ++i) {
      c->ref.N[q][i] = N[q * nbf + i];
      for (int d = 0; d < rdim; ++d) c->ref.dN[q][i][d] = dNdxi[(q * nbf + i) * rdim + d];
    }
    for (int j = 0; j < ngeo; ++j)
      for (int d = 0; d < rdim; ++d) c->ref.dM[q][j][d] = dMdxi[(q * ngeo + j) * rdim + d];
  }
  c->rdim = rdim; c->nqp = nqp; c->nbf = nbf; c->ngeo = ngeo; c->vdim = vdim;
  c->have_cv = true;
  c->tab_dirty = true;
  ++c->aux_version;
  return FEM_OK;
}

int32_t fem_cellparam_upload(fem_ctx* c, int64_t ncells, const double* param) {
  if (!c || !param || ncells != c->ncells) return set_error(c, FEM_ERR_ARG, "fem_cellparam_upload: length must equal ncells");
  if (c->ncolors == 0) return set_error(c, FEM_ERR_STATE, "upload the coloring before cell parameters");
  int rc;
  if ((rc = dev_alloc(c, &c->d_param_cell, ncells))) return rc;
  if ((rc = dev_alloc(c, &c->d_cparam, ncells))) return rc;
  FEM_HIP(c, hipMemcpyAsync(c->d_param_cell, param, sizeof(double) * ncells, hipMemcpyHostToDevice, c->stream));
  k_gather_param<<<grid_for(ncells), 256, 0, c->stream>>>(ncells, c->d_ccell, c->d_param_cell, c->d_cparam);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  c->have_param = true;
  ++c->aux_version;
  return FEM_OK;
}

int32_t fem_matrix_download(fem_ctx* c, double* nzval, double* f) {
  if (!c || c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no matrix");
  if (nzval) FEM_HIP(c, hipMemcpyAsync(nzval, c->d_nzval, sizeof(double) * c->nnz, hipMemcpyDeviceToHost, c->stream));
  if (f) FEM_HIP(c, hipMemcpyAsync(f, c->d_f, sizeof(double) * c->ndofs, hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  return FEM_OK;
}

int32_t fem_matrix_upload(fem_ctx* c, const double* nzval, const double* f) {
  if (!c || c->nnz == 0) return set_error(c, FEM_ERR_STATE, "no matrix");
  if (nzval) FEM_HIP(c, hipMemcpyAsync(c->d_nzval, nzval, sizeof(double) * c->nnz, hipMemcpyHostToDevice, c->stream));
  if (f) FEM_HIP(c, hipMemcpyAsync(c->d_f, f, sizeof(double) * c->ndofs, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  if (nzval) ++c->nzval_version;
  return FEM_OK;
}

int32_t fem_solution_download(fem_ctx* c, double* u) {
  if (!c || !u || c->ndofs == 0) return set_error(c, FEM_ERR_STATE, "no solution");
  FEM_HIP(c, hipMemcpyAsync(u, c->d_u, sizeof(double) * c->ndofs, hipMemcpyDeviceToHost, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  return FEM_OK;
}

int32_t fem_solution_upload(fem_ctx* c, const double* u) {
  if (!c || !u || c->ndofs == 0) return set_error(c, FEM_ERR_STATE, "no dofs");
  FEM_HIP(c, hipMemcpyAsync(c->d_u, u, sizeof(double) * c->ndofs, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  return FEM_OK;
}

int32_t fem_device_pointer(fem_ctx* c, int32_t which, void** ptr, int64_t* nelem) {
  if (!c || !ptr) return FEM_ERR_ARG;
  switch (which) {
    case 0:   // writable: the solver's operator copies are rebuilt from nzval at the next solve
      *ptr = c->d_nzval;
      if (nelem) *nelem = c->nnz;
      ++c->nzval_version;
      break;
    case 1: *ptr = c->d_f; if (nelem) *nelem = c->ndofs; break;
    case 2: *ptr = c->d_u; if (nelem) *nelem = c->ndofs; break;
    default: return set_error(c, FEM_ERR_ARG, "unknown array");
  }
  return FEM_OK;
}

int32_t fem_get_timings(const fem_ctx* c, fem_timings* t) {
  if (!c || !t) return FEM_ERR_ARG;
  *t = c->timings;
  return FEM_OK;
}

}  // extern "C"
