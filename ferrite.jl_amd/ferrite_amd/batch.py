"""Batched assembly + solve of many independent scalar Q1 Poisson problems.

Each problem is set up exactly as the serial loop of
docs/src/literate-tutorials/heat_equation.jl:186-218 sets up one: its own Grid
(generate_grid), DofHandler with one Lagrange{1} field, create_sparsity_pattern,
create_coloring and a ConstraintHandler with a homogeneous Dirichlet condition.
The batch then runs every problem's start_assemble / assemble! / apply! / K \\ f
on the device at once (include/ferrite_mi355x_batch.h):

    batch = BatchedPoisson()
    for grid, dh, colors, ch, k in problems:
        batch.add(grid, dh, colors, ch, k)
    batch.finalize()
    batch.autotune("mixed")
    res = batch.run("mixed", rtol=1e-8)
    nzval, f, u = batch.download(0)
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .constraints import ConstraintHandler, Dirichlet
from .dofs import DofHandler, create_coloring, create_sparsity_pattern
from .mesh import Hexahedron, Quadrilateral, generate_grid
from .refel import Lagrange, QuadratureRule
from .values import CellValues


class BatchResult:
    def __init__(self, info, completion_ms, iterations):
        self.assemble_ms = info.assemble_ms
        self.apply_ms = info.apply_ms
        self.solve_ms = info.solve_ms
        self.total_ms = info.total_ms
        self.converged = info.converged
        self.total_iterations = info.total_iterations
        self.completion_ms = completion_ms
        self.iterations = iterations

    def percentile(self, q):
        return float(np.percentile(self.completion_ms, q))


class BatchedPoisson:
    """Many Poisson problems -K∇²u = f (per-cell k) in one device context."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = L.lib().fem_batch_create(int(device), C.byref(h))
        L.raise_for(rc, what="fem_batch_create failed (is a GPU visible?)")
        self._h = h
        self.problems = []
        self.finalized = False
        self.source = None

    def _call(self, name, *args):
        rc = getattr(L.lib(), name)(self._h, *args)
        L.raise_for(rc, self._h, name, batch=True)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            L.lib().fem_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, grid, dh, colors, ch, k, source=1.0):
        """Add one problem: its Grid, closed DofHandler, create_coloring colours, closed
        ConstraintHandler (homogeneous), per-cell conductivity k[ncells] and source f."""
        if self.finalized:
            raise RuntimeError("the batch is finalized")
        if grid.celltype not in (Quadrilateral, Hexahedron):
            raise ValueError("batched problems are Quadrilateral or Hexahedron Q1 grids")
        if len(dh.field_names) != 1 or dh.ndofs_per_cell != grid.cells.shape[1]:
            raise ValueError("batched problems have one scalar Lagrange{1} field")
        if np.any(np.asarray(ch.inhomogeneities) != 0.0):
            raise ValueError("batched problems take homogeneous Dirichlet conditions")
        colptr, rowval = create_sparsity_pattern(dh)
        sizes = np.array([len(c) for c in colors], dtype=np.int64)
        cptr = np.zeros(len(colors) + 1, dtype=np.int64)
        cptr[1:] = np.cumsum(sizes)
        ccells = L.as_i64(np.concatenate([np.asarray(c, dtype=np.int64) for c in colors]))
        xyz = L.as_f64(grid.nodes)
        cells = L.as_i64(grid.cells)
        cdofs = L.as_i64(dh.cell_dofs)
        kk = L.as_f64(k)
        if kk.shape != (grid.ncells,):
            raise ValueError("k must hold one value per cell")
        pres = L.as_i64(ch.prescribed_dofs)
        colptr, rowval = L.as_i64(colptr), L.as_i64(rowval)
        self._call("fem_batch_add_problem", grid.sdim, grid.nnodes, L.ptr(xyz, C.c_double), grid.ncells,
                   L.ptr(cells, C.c_int64), dh.ndofs, L.ptr(cdofs, C.c_int64), L.ptr(colptr, C.c_int64),
                   L.ptr(rowval, C.c_int64), len(colors), L.ptr(cptr, C.c_int64), L.ptr(ccells, C.c_int64),
                   L.ptr(kk, C.c_double), len(pres), L.ptr(pres, C.c_int64), float(source))
        self.problems.append(dict(ndofs=dh.ndofs, nnz=len(rowval), ncells=grid.ncells, dim=grid.sdim,
                                  colptr=colptr, rowval=rowval))
        self.source = source
        return len(self.problems) - 1

    def finalize(self):
        tabs = []
        for shape in ("RefQuadrilateral", "RefHexahedron"):
            cv = CellValues(QuadratureRule(shape, 2), Lagrange(shape, 1))
            _, _, _, _, _, N, dN, _, w = cv.device_tables()
            tabs += [L.as_f64(N), L.as_f64(dN), L.as_f64(w)]
        self._tabs = tabs
        self._call("fem_batch_finalize", *[L.ptr(t, C.c_double) for t in tabs])
        self.finalized = True
        return self

    def autotune(self, precision="mixed"):
        """Times every (cells per workgroup, chunk) launch shape on this batch; keeps the fastest."""
        cpw, chunk = C.c_int32(0), C.c_int32(0)
        self._call("fem_batch_autotune", _prec(precision), C.byref(cpw), C.byref(chunk))
        self.shape = (cpw.value, chunk.value)
        return self.shape

    def assemble(self, precision="mixed", cells_per_wg=0, chunk=0):
        ms = C.c_double(0.0)
        self._call("fem_batch_assemble_only", _prec(precision), int(cells_per_wg), int(chunk), C.byref(ms))
        return ms.value

    def run(self, precision="mixed", rtol=1e-8, maxit=0, cells_per_wg=0, chunk=0, ru_delta=0.0):
        """Assemble, apply! and solve every problem (the solver runs on the sliced copy of K)."""
        cfg = L.BatchCfg(precision=_prec(precision), cells_per_wg=int(cells_per_wg), chunk=int(chunk),
                         maxit=int(maxit), rtol=float(rtol), ru_delta=float(ru_delta))
        info = L.BatchInfo()
        rc = L.lib().fem_batch_run(self._h, C.byref(cfg), C.byref(info))
        n = len(self.problems)
        ms = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        L.lib().fem_batch_problem_stats(self._h, L.ptr(ms, C.c_double), L.ptr(it, C.c_int32))
        L.raise_for(rc, self._h, "fem_batch_run", batch=True)
        return BatchResult(info, ms, it)

    def download(self, i):
        p = self.problems[i]
        nz = np.zeros(p["nnz"])
        f = np.zeros(p["ndofs"])
        u = np.zeros(p["ndofs"])
        self._call("fem_batch_download", int(i), L.ptr(nz, C.c_double), L.ptr(f, C.c_double), L.ptr(u, C.c_double))
        return nz, f, u

    @property
    def ncells(self):
        return sum(p["ncells"] for p in self.problems)

    @property
    def ndofs(self):
        return sum(p["ndofs"] for p in self.problems)


def _prec(precision):
    if precision in ("mixed", L.FEM_MIXED):
        return L.FEM_MIXED
    if precision in ("fp64", L.FEM_FP64):
        return L.FEM_FP64
    raise ValueError(f"unknown precision {precision!r}")


def batch_sizes(nproblems=64, seed=0, lo=2e3, hi=2e5):
    """(celltype, nel) per problem: 2D/3D 50/50 by seed, cell counts log-uniform in [lo, hi],
    grid dimensions derived from the count (near-square / near-cube)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nproblems):
        dim = 2 if rng.random() < 0.5 else 3
        target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if dim == 2:
            nx = max(2, int(round(np.sqrt(target))))
            ny = max(2, int(round(target / nx)))
            out.append((Quadrilateral, (nx, ny)))
        else:
            nx = max(2, int(round(target ** (1.0 / 3.0))))
            ny = max(2, int(round(np.sqrt(target / nx))))
            nz = max(2, int(round(target / (nx * ny))))
            out.append((Hexahedron, (nx, ny, nz)))
    return out


def conductivity(ncells, rng, sigma=0.5):
    return np.exp(rng.normal(0.0, sigma, size=ncells))


def poisson_problem(celltype, nel, k, coloring="workstream"):
    """One problem of the batch set up like heat_equation.jl: Dirichlet u = 0 on every boundary facet."""
    grid = generate_grid(celltype, nel)
    dh = DofHandler(grid)
    dh.add("u", Lagrange(grid.refshape, 1))
    dh.close()
    facets = []
    for name in sorted(grid.facetsets):
        facets.extend(grid.facetsets[name])
    ch = ConstraintHandler(dh)
    ch.add(Dirichlet("u", facets, lambda x, t: 0.0))
    ch.close()
    colors = create_coloring(grid, coloring)
    return grid, dh, colors, ch, np.asarray(k, dtype=np.float64)


def make_batch(nproblems=64, seed=0, lo=2e3, hi=2e5, coloring="workstream", device=0, sizes=None):
    """Build and finalize the batch of `nproblems` seeded problems; returns (batch, specs)."""
    rng = np.random.default_rng(seed + 1)
    specs = sizes if sizes is not None else batch_sizes(nproblems, seed, lo, hi)
    batch = BatchedPoisson(device)
    setups = []
    for celltype, nel in specs:
        k = conductivity(int(np.prod(nel)), rng)
        grid, dh, colors, ch, k = poisson_problem(celltype, nel, k, coloring)
        batch.add(grid, dh, colors, ch, k)
        setups.append((celltype, nel, k))
    batch.finalize()
    return batch, setups
