// Linear elasticity element kernels (vector Lagrange) and load entry points.
#include "fem_internal.h"

namespace fem {

int assemble_elasticity(fem_ctx* c, const fem_assemble_cfg* cfg) {
  (void)cfg;
  return set_error(c, FEM_ERR_ARG, "elasticity kernels are not available in this build");
}

__global__ void k_point_load(int64_t n, const int32_t* __restrict__ dofs, const double* __restrict__ v,
                             double* __restrict__ f) {
  // serial in one thread: loads may repeat a dof, and their order must be the caller's
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = 0; i < n; ++i) f[dofs[i]] += v[i];
}

}  // namespace fem

using namespace fem;

extern "C" {

int32_t fem_point_load(fem_ctx* c, int64_t n, const int64_t* dofs, const double* values) {
  if (!c || n < 0 || (n > 0 && (!dofs || !values)) || c->ndofs == 0)
    return set_error(c, FEM_ERR_ARG, "fem_point_load: invalid arguments");
  if (n == 0) return FEM_OK;
  std::vector<int32_t> d(n);
  for (int64_t i = 0; i < n; ++i) {
    if (dofs[i] < 1 || dofs[i] > c->ndofs) return set_error(c, FEM_ERR_ARG, "point load dof out of range");
    d[i] = (int32_t)(dofs[i] - 1);
  }
  int32_t* dd = nullptr;
  double* dv = nullptr;
  int rc;
  if ((rc = dev_alloc(c, &dd, n)) || (rc = dev_alloc(c, &dv, n))) { dev_free(&dd); return rc; }
  FEM_HIP(c, hipMemcpyAsync(dd, d.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  FEM_HIP(c, hipMemcpyAsync(dv, values, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  k_point_load<<<1, 64, 0, c->stream>>>(n, dd, dv, c->d_f);
  FEM_HIP(c, hipGetLastError());
  FEM_HIP(c, hipStreamSynchronize(c->stream));
  dev_free(&dd);
  dev_free(&dv);
  return FEM_OK;
}

int32_t fem_facet_traction(fem_ctx* c, int64_t, const int64_t*, int32_t, int32_t, const double*, const double*,
                           const double*, const double*) {
  return set_error(c, FEM_ERR_STATE, "facet traction is not available in this build");
}

int32_t fem_adjoint_compliance(fem_ctx* c, const double*, double, double, double, double*, double*, double*) {
  return set_error(c, FEM_ERR_STATE, "the compliance adjoint is not available in this build");
}

}  // extern "C"
