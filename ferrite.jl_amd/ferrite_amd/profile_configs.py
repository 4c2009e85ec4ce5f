"""Full-size runs of the heat, elasticity, differentiation and batching workloads (one JSON line each).

    python -m ferrite_amd.profile_configs heat         [--nx 100 --ny 100]
    python -m ferrite_amd.profile_configs elasticity   [--n 48]
    python -m ferrite_amd.profile_configs cantilever   [--nx 512 --ny 256 --fd 64]
    python -m ferrite_amd.profile_configs batch        [--problems 64]

heat        The heat equation of heat_equation.jl on generate_grid(Quadrilateral, (nx, ny)) over
            [-1,1]^2, Q1, 2x2 Gauss, k = 1, f = 1, u = 0 on all four boundary facet sets: assembly
            kernel time (best of `reps`), apply! and the FP64 CG solve to 1e-10.
elasticity  Q2 hex^3 small-strain elasticity on generate_grid(Hexahedron, (n, n, n)) over [0,1]^3,
            3x3x3 Gauss, nu = 0.3, E_e ~ U[0.5, 2.0] (seed 0), clamped 'left', traction (0,0,-1)
            on 'right' (3x3 facet rule).  Times the FP64 element path and the FP32 MFMA element
            path, each through the owner-computes path (element phase and gather phase timed
            apart with HIP events) and through the colored scatter; reports K differences
            between them, the element contraction rate (dense B^T D B count, 2,283,228
            flop/cell, and the MFMA work issued), the gather's bytes and GB/s, and the u error
            of the FP32 element path split into element rounding (backward error of the FP32
            K at the FP64 solution) and CG tolerance (both solved to rtol and to 5e-11).
cantilever  Plane-strain Q1^2 cantilever on (nx, ny) cells over [0,2]x[0,1], E = E_min +
            theta^3 (1 - E_min), theta ~ U[0.2, 1.0] (seed 0), clamped at x = 0, point load
            (0, -1) at (2, 0.5).  Times J (device CG) and the adjoint kernel, then checks `fd`
            entries of dJ/dtheta and dJ/dX (random cells / node coordinates whose supports share
            no node) against central differences at h = 1e-6 and 1e-4 that carry no solve
            error: K_ff is factored once on the host and each J(theta +- h) - J(theta) follows
            from a Woodbury update (exact_central_differences); the adjoint kernel runs on the
            LU solution.
batch       `problems` seeded scalar Poisson problems (2D/3D, log-uniform cell counts in
            [2e3, 2e5]) assembled and solved (mixed CG to 1e-8) in one batched context after
            the launch-shape autotuner; total cells/s and p50 / p99 completion times.

Each function takes an optional `cpu` callable; when given, it is called once after the device
measurements with the host-side inputs of the workload (keyword arguments) and its result is
stored as `cpu_baseline` (the bench passes the timed CPU restatement of the same workload).
"""
import argparse
import json
import os
import sys
import time

DENSE_BTDB_FLOP_PER_CELL = 2_283_228          # 27 q x 2 x (6*6*81 + 81*6*81)
# the element kernels (FP32 and FP64 MFMA): nine products, each 4 tiles of 16x16 over 7 steps of
# v_mfma_{f32,f64}_16x16x4
MFMA_ISSUED_FLOP_PER_CELL = 9 * 4 * 7 * 16 * 16 * 4 * 2
MFMA_FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# tight CG tolerance of the C3 error split (FP64 Jacobi CG on the 48^3 system stagnates
# between 1e-11 and 4e-11, depending on the rounding of K)
TIGHT_RTOL = 5e-11
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def q2_gather_bytes(ncells, nnode, nnz, ke_bytes):
    """Algorithmic bytes of the owner-computes Q2 gather (phase 2): every cell's stored element
    matrix read once (27 blocks of 252 values), its 27 x 27 rank bytes and 27 list entries,
    two colptr values and the list bounds per dof node, and K.nzval written once."""
    return ncells * (27 * 252 * ke_bytes + 27 * 27 + 27 * 4) + nnode * (8 + 4) + nnz * 8


def elasticity(args, cpu=None):
    import numpy as np
    import scipy.sparse as sp
    import ferrite_amd as fa
    n = args.n
    t0 = time.perf_counter()
    g = fa.generate_grid("Hexahedron", (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ip = fa.Lagrange("RefHexahedron", 2) ** 3
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 3), ip)
    t_host = time.perf_counter()
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    t_host = time.perf_counter() - t_host
    t_dev = time.perf_counter()
    dh_d = fa.DofHandler(g).add("u", ip).close(device=0)
    K_d = fa.allocate_matrix(fa.MI355XCSC, dh_d)
    t_dev = time.perf_counter() - t_dev
    same = bool(np.array_equal(dh_d.cell_dofs, dh.cell_dofs) and np.array_equal(K_d.colptr, K.colptr)
                and np.array_equal(K_d.rowval, K.rowval))
    log(f"setup: host {t_host:.2f} s, device {t_dev:.2f} s, identical {same}")
    del K_d, dh_d
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    E = np.random.default_rng(args.seed).uniform(0.5, 2.0, g.ncells)
    el = fa.ElasticityElement(E=E, nu=0.3)
    fv = fa.FacetValues(fa.FacetQuadratureRule("RefHexahedron", 3), ip)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0, 0.0])).close()
    log(f"setup {time.perf_counter() - t0:.1f}s: {g.ncells} cells, {dh.ndofs} dofs, nnz {K.nnz}, "
        f"{len(colors)} colors")
    nc, nnode = g.ncells, dh.ndofs // 3
    out = {"workload": f"Q2 hex^3 elasticity {n}^3, 3x3x3 Gauss, E ~ U[0.5,2], nu 0.3",
           "cells": nc, "dofs": dh.ndofs, "nnz": K.nnz, "colors": len(colors),
           "setup": {"host_s": t_host, "device_s": t_dev, "device_equals_host": same},
           "units": {"dense_btdb_flop_per_cell": DENSE_BTDB_FLOP_PER_CELL,
                     "mfma_issued_flop_per_cell": MFMA_ISSUED_FLOP_PER_CELL}}
    nzs = {}
    for prec in ("fp64", "mixed"):
        for asm in ("gather", "scatter"):
            fa.set_tuning(K, asm_mode=0 if asm == "gather" else 1)
            runs = []
            for _ in range(args.reps):
                fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
                runs.append(fa.timings(K))
            best = min(runs, key=lambda t: t["assemble_kernel_ms"])
            ms = best["assemble_kernel_ms"]
            rec = {"assembly_ms": ms, "assembly_ms_all": [t["assemble_kernel_ms"] for t in runs],
                   "cells_per_s": nc / (ms * 1e-3),
                   "dense_btdb_tflops": DENSE_BTDB_FLOP_PER_CELL * nc / (ms * 1e-3) / 1e12}
            if asm == "gather":
                e_ms, s_ms = best["element_ms"], best["scatter_ms"]
                kb = 4 if prec == "mixed" else 8
                gb = q2_gather_bytes(nc, nnode, K.nnz, kb)
                rec.update({"element_ms": e_ms, "scatter_ms": s_ms,
                            "element_dense_btdb_tflops": DENSE_BTDB_FLOP_PER_CELL * nc / (e_ms * 1e-3) / 1e12,
                            "element_ke_write_GBs": nc * 6804 * kb / (e_ms * 1e-3) / 1e9,
                            "scatter_bytes": gb, "scatter_GBs": gb / (s_ms * 1e-3) / 1e9,
                            "scatter_frac_hbm": gb / (s_ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
                # both element kernels issue the same MFMA work (FP32 or FP64 16x16x4)
                peak = MFMA_FP32_PEAK_TFLOPS if prec == "mixed" else FP64_PEAK_TFLOPS
                rec["element_mfma_issued_tflops"] = MFMA_ISSUED_FLOP_PER_CELL * nc / (e_ms * 1e-3) / 1e12
                rec["element_mfma_frac_issued"] = rec["element_mfma_issued_tflops"] / peak
                log(f"{prec} {asm}: assembly {ms:.3f} ms (element {e_ms:.3f} ms, gather {s_ms:.3f} ms "
                    f"= {rec['scatter_GBs']:.0f} GB/s)")
            else:
                log(f"{prec} {asm}: assembly {ms:.3f} ms")
            out[f"{prec}_{asm}"] = rec
            nzs[(prec, asm)] = K.nzval()
    for prec in ("fp64", "mixed"):
        a, b = nzs[(prec, "scatter")], nzs[(prec, "gather")]
        out[f"{prec}_K_rel_diff_scatter_vs_gather"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    del nzs[("fp64", "scatter")], nzs[("mixed", "scatter")]
    fa.set_tuning(K, asm_mode=0)
    a, b = nzs[("mixed", "gather")], nzs[("fp64", "gather")]
    out["K_rel_err_mixed_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    # u error of the FP32 element path, split into the element rounding and the CG tolerance:
    # both systems solved to args.rtol and to TIGHT_RTOL, and the backward error of the FP32-element
    # matrix at the FP64 solution, ||(K32 - K64) u64|| / ||f||, which is what the rounding of Ke
    # alone contributes to the residual
    sols, kc = {}, {}
    for prec in ("fp64", "mixed"):
        fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
        fa.apply_traction(K, f, fv, g.facetsets["right"], (0.0, 0.0, -1.0))
        fa.apply(K, f, ch)
        kc[prec] = K.nzval()
        fvec = f.to_numpy()
        for rtol in (args.rtol, TIGHT_RTOL):
            t1 = time.perf_counter()
            res = fa.solve(K, f, mode="cg", rtol=rtol, maxit=20000)
            ts = time.perf_counter() - t1
            sols[(prec, rtol)] = res.u
            out[f"{prec}_gather"][f"solve_rtol_{rtol:g}"] = {"solve_s": ts, "cg_iterations": res.iterations,
                                                             "relres": res.relres}
            log(f"{prec}: CG to {rtol:g}: {res.iterations} its {ts:.2f}s relres {res.relres:.2e}")
    for rtol in (args.rtol, TIGHT_RTOL):
        u32, u64 = sols[("mixed", rtol)], sols[("fp64", rtol)]
        out[f"u_rel_err_mixed_vs_fp64_rtol_{rtol:g}"] = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    u64 = sols[("fp64", TIGHT_RTOL)]
    cp = np.asarray(K.colptr, dtype=np.int64) - 1
    rv = np.asarray(K.rowval, dtype=np.int64) - 1
    D = sp.csc_matrix((kc["mixed"] - kc["fp64"], rv, cp), shape=(dh.ndofs, dh.ndofs))
    out["backward_err_fp32_Ke_at_u64"] = float(np.linalg.norm(D @ u64) / np.linalg.norm(fvec))
    out["K_rel_err_mixed_vs_fp64_constrained"] = float(np.linalg.norm(kc["mixed"] - kc["fp64"]) /
                                                       np.linalg.norm(kc["fp64"]))
    log(f"u error mixed vs fp64: {out[f'u_rel_err_mixed_vs_fp64_rtol_{args.rtol:g}']:.2e} at rtol {args.rtol:g}, "
        f"{out[f'u_rel_err_mixed_vs_fp64_rtol_{TIGHT_RTOL:g}']:.2e} at {TIGHT_RTOL:g}; backward error of the FP32 Ke "
        f"{out['backward_err_fp32_Ke_at_u64']:.2e}")
    if cpu is not None:
        del D, kc, sols, nzs
        out["cpu_baseline"] = cpu(g=g, dh=dh, E=E, nu=0.3, colors=colors, colptr=K.colptr, rowval=K.rowval)
        cb = out["cpu_baseline"]
        if cb:
            out["fp64_gather"]["speedup_vs_cpu_assembly"] = out["fp64_gather"]["cells_per_s"] / cb["value"]
    return out


def _disjoint_picks(rng, candidates, node_sets, want):
    """Up to `want` of `candidates` (in the random order of rng) whose node sets are pairwise disjoint."""
    used, out = set(), []
    for k in rng.permutation(candidates):
        s = node_sets(int(k))
        if used.isdisjoint(s):
            used.update(s)
            out.append(int(k))
            if len(out) == want:
                break
    return out


def pick_fd_entries(g, rng, n_cells, n_coords):
    """Random cells, and random (node, direction) coordinates, for exact_central_differences:
    no two chosen cells share a node, and no two chosen nodes' cell patches share a node."""
    import numpy as np
    cells0 = g.cells - 1
    node_cells = {}
    for e, row in enumerate(cells0):
        for nd in row:
            node_cells.setdefault(int(nd), []).append(e)
    cells = _disjoint_picks(rng, np.arange(g.ncells), lambda e: set(cells0[e].tolist()), n_cells)
    nodes = _disjoint_picks(rng, np.arange(g.nnodes), lambda nd: set(cells0[node_cells[nd]].ravel().tolist()),
                            n_coords)
    return cells, [(nd, int(rng.integers(2))) for nd in nodes]


def exact_central_differences(prob, theta, cells, coords, hs, chunk=64):
    """Central differences of J(θ, X) = fᵀu, K u = f, that carry no solve error.

    The free-dof block K_ff of the assembled, constrained system is factored once (sparse LU
    on the host, two steps of iterative refinement).  Changing θ_e or one node coordinate
    changes K_ff only on the free dofs U of the cells involved, K_ff + U M Uᵀ, and by the
    Woodbury identity
        J(M) − J = −vᵀ M (I + G M)⁻¹ v,   v = Uᵀ u,   G = Uᵀ K_ff⁻¹ U,
    which is formed directly instead of as the difference of two J values.  M comes from device
    assemblies: Ke⁰ of the chosen cells from one assembly with E = their indicator, and ΔK of a
    node coordinate as K(X ± h) − K(X) from assemblies with every chosen coordinate moved at
    once.  The chosen cells, and the cell patches of the chosen nodes, must share no node (so the
    blocks do not overlap); `coords` holds (node index, direction) pairs, 0-based.

    Returns u (full vector, the LU solution), the refinement residual ‖f − K u‖/‖f‖ and
    {h: (fd_theta, fd_x)}.  Leaves the device mesh at X and K, f assembled at θ."""
    import numpy as np
    import scipy.sparse.linalg as sla
    from . import device as D
    from .autodiff import _mesh_args
    K, f, ch = prob.K, prob.f, prob.ch
    g, dh = K.dh.grid, K.dh
    n = dh.ndofs
    E = D.simp_modulus(theta, prob.emin, prob.penal)

    def assemble(Evals, X=None, constrain=False):
        if X is not None:
            g.nodes[...] = X
            K.ctx.call("fem_mesh_upload", *_mesh_args(K))
            K._colors = None
        D.assemble_global(D.start_assemble(K, f), prob.cv, prob.colors, D.ElasticityElement(E=Evals, nu=prob.nu))
        if constrain:
            prob.loads(K, f)
            D.apply(K, f, ch)
        return K.to_scipy()

    X0 = g.nodes.copy()
    free = np.ones(n, dtype=bool)
    free[np.asarray(ch.prescribed_dofs, dtype=np.int64) - 1] = False
    fidx = np.flatnonzero(free)
    pos = np.full(n, -1, dtype=np.int64)
    pos[fidx] = np.arange(len(fidx))

    # blocks first (they overwrite K and f), the constrained system last
    ind = np.zeros(g.ncells)
    ind[cells] = 1.0
    K0 = assemble(ind)
    Kx = {}
    for h in hs:
        for sgn in (1, -1):
            Xs = X0.copy()
            for nd, d in coords:
                Xs[nd, d] += sgn * h
            Kx[(h, sgn)] = assemble(E, Xs)
    Kraw = assemble(E, X0)

    cell_dofs = dh.cell_dofs - 1
    node_cells = {}
    for c_, row in enumerate(g.cells - 1):
        for nd in row:
            node_cells.setdefault(int(nd), []).append(c_)
    Ut = [np.unique(cell_dofs[e]) for e in cells]
    Ux = [np.unique(cell_dofs[node_cells[nd]]) for nd, _ in coords]
    # a block whose dofs are all free is annihilated by rigid translations (per component)
    whole_t = [bool(free[u_].all()) for u_ in Ut]
    whole_x = [bool(free[u_].all()) for u_ in Ux]
    Ut = [u_[free[u_]] for u_ in Ut]
    Ux = [u_[free[u_]] for u_ in Ux]
    comp = np.zeros(n, dtype=np.int64)
    for d in range(g.sdim):
        comp[cell_dofs[:, d::g.sdim]] = d

    A = assemble(E, None, constrain=True)
    Aff = A[fidx][:, fidx].tocsc()
    fvec = D.DeviceVector(K, 1).to_numpy()
    ff = fvec[fidx]
    t0 = time.perf_counter()
    lu = sla.splu(Aff, permc_spec="MMD_AT_PLUS_A", options=dict(SymmetricMode=True))
    log(f"sparse LU of K_ff ({len(fidx)} dofs): {time.perf_counter() - t0:.1f}s, fill {lu.L.nnz + lu.U.nnz}")
    uf = lu.solve(ff)
    for _ in range(2):
        uf += lu.solve(ff - Aff @ uf)
    refine_res = float(np.linalg.norm(ff - Aff @ uf) / np.linalg.norm(ff))
    u = np.zeros(n)
    u[fidx] = uf

    # G = Uᵀ K_ff⁻¹ U for every block, from one set of unit right-hand sides
    allU = np.unique(np.concatenate(Ut + Ux))
    col = {int(d): i for i, d in enumerate(allU)}
    Z = np.empty((len(fidx), len(allU)))
    t0 = time.perf_counter()
    for s in range(0, len(allU), chunk):
        B = np.zeros((len(fidx), min(chunk, len(allU) - s)))
        B[pos[allU[s:s + B.shape[1]]], np.arange(B.shape[1])] = 1.0
        Z[:, s:s + B.shape[1]] = lu.solve(B)
        log(f"K_ff^-1 columns {s + B.shape[1]}/{len(allU)}: {time.perf_counter() - t0:.1f}s")

    def dJ(U, M, shift):
        # With t a translation of the block (M t = 0, so (I + G M)⁻¹ t = t), vᵀM(I + GM)⁻¹v equals
        # wᵀM(I + GM)⁻¹w for w = v − t.  Taking t = the per-component mean of v removes the bulk
        # displacement, which would otherwise multiply the rounding of M (entries ~E, ~eps E
        # error, not translation-free) by |v|².
        p = pos[U]
        G = Z[p][:, [col[int(d)] for d in U]]
        v = u[U].copy()
        if shift:
            for d in range(g.sdim):
                m = comp[U] == d
                v[m] -= v[m].mean()
        return -float(v @ M @ np.linalg.solve(np.eye(len(U)) + G @ M, v))

    def dense(S, U):
        return S[:, U][U].toarray()

    out = {}
    for h in hs:
        fdt = []
        for e, U, sh in zip(cells, Ut, whole_t):
            K0e = dense(K0, U)
            Ep = D.simp_modulus(theta[e] + h, prob.emin, prob.penal) - E[e]
            Em = D.simp_modulus(theta[e] - h, prob.emin, prob.penal) - E[e]
            fdt.append((dJ(U, Ep * K0e, sh) - dJ(U, Em * K0e, sh)) / (2 * h))
        fdx = []
        for U, sh in zip(Ux, whole_x):
            base = dense(Kraw, U)
            fdx.append((dJ(U, dense(Kx[(h, 1)], U) - base, sh) - dJ(U, dense(Kx[(h, -1)], U) - base, sh)) / (2 * h))
        out[h] = (np.array(fdt), np.array(fdx))
    return u, refine_res, out


def cantilever(args, cpu=None):
    import numpy as np
    import ferrite_amd as fa
    from ferrite_amd import _lib as L
    from ferrite_amd.autodiff import ComplianceProblem
    nx, ny = args.nx, args.ny
    t0 = time.perf_counter()
    g = fa.generate_grid("Quadrilateral", (nx, ny), (0.0, 0.0), (2.0, 1.0))
    ip = fa.Lagrange("RefQuadrilateral", 1) ** 2
    cv = fa.CellValues(fa.QuadratureRule("RefQuadrilateral", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    rng = np.random.default_rng(args.seed)
    theta = rng.uniform(0.2, 1.0, g.ncells)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0])).close()
    tip = int(np.nonzero(np.all(np.isclose(g.nodes, [2.0, 0.5]), axis=1))[0][0]) + 1
    c = int(np.nonzero(np.any(g.cells == tip, axis=1))[0][0])
    a = int(np.nonzero(g.cells[c] == tip)[0][0])
    ydof = int(dh.cell_dofs[c, 2 * a + 1])

    def loads(K_, f_):
        fa.add_point_load(K_, f_, [ydof], [-1.0])
    prob = ComplianceProblem(K, f, cv, colors, ch, loads, rtol=args.rtol)
    log(f"setup {time.perf_counter() - t0:.1f}s: {g.ncells} cells, {dh.ndofs} dofs, tip node {tip}")
    prob.compliance(theta)      # solver set-up and graph capture
    t1 = time.perf_counter()
    J, dth_cg, dx_cg, res = prob.compliance(theta)
    tj = time.perf_counter() - t1
    t1 = time.perf_counter()
    fa.compliance_gradient(K, theta)
    tg = time.perf_counter() - t1
    log(f"J = {J:.10e}: forward {tj:.2f}s ({res.iterations} CG its), adjoint kernel {tg * 1e3:.2f} ms")

    cells, coords = pick_fd_entries(g, np.random.default_rng(args.seed + 1), args.fd // 2, args.fd - args.fd // 2)
    hs = (1e-6, 2e-6, 1e-4)
    t1 = time.perf_counter()
    u, refine_res, fd = exact_central_differences(prob, theta, cells, coords, hs)
    tfd = time.perf_counter() - t1
    K.ctx.call("fem_solution_upload", L.ptr(u, L.C.c_double))
    J_lu, dth, dx = fa.compliance_gradient(K, theta)
    ad_t = dth[cells]
    ad_x = np.array([dx[nd, d] for nd, d in coords])

    def stats(fdv, ad):
        rel = np.abs(fdv - ad) / np.abs(ad)
        return {"entries": len(ad), "normwise_rel_err": float(np.linalg.norm(fdv - ad) / np.linalg.norm(ad)),
                "max_entry_rel_err": float(rel.max()), "median_entry_rel_err": float(np.median(rel)),
                "entries_above_1e-6": int(np.sum(rel > 1e-6)),
                "min_abs_entry": float(np.abs(ad).min()), "max_abs_entry": float(np.abs(ad).max())}
    out = {"workload": f"plane-strain Q1^2 cantilever {nx}x{ny}, SIMP p=3, E_min 1e-3, point load at node {tip}",
           "cells": g.ncells, "dofs": dh.ndofs, "J": J, "forward_s": tj, "cg_iterations": res.iterations,
           "cg_rtol": args.rtol, "relres": res.relres, "adjoint_kernel_ms": tg * 1e3,
           "fd_method": "exact central differences: one sparse LU of K_ff (host), Woodbury update per "
                        "entry with M from device assemblies; adjoint kernel evaluated on the LU solution",
           "fd_entries": {"cells": [int(e) for e in cells], "coords": [[int(nd), int(d)] for nd, d in coords]},
           "lu_refined_relres": refine_res, "fd_wall_s": tfd, "J_lu": J_lu,
           "J_rel_diff_cg_vs_lu": abs(J - J_lu) / abs(J_lu),
           "grad_rel_diff_cg_vs_lu": {
               "dJdtheta": float(np.linalg.norm(dth_cg - dth) / np.linalg.norm(dth)),
               "dJdX": float(np.linalg.norm(dx_cg - dx) / np.linalg.norm(dx))}}
    for h in hs:
        fdt, fdx = fd[h]
        out[f"fd_h{h:g}"] = {"dJdtheta": stats(fdt, ad_t), "dJdX": stats(fdx, ad_x)}
        log(f"h = {h:g}: dJ/dtheta max rel {out[f'fd_h{h:g}']['dJdtheta']['max_entry_rel_err']:.2e}, "
            f"dJ/dX max rel {out[f'fd_h{h:g}']['dJdX']['max_entry_rel_err']:.2e}")
    (t1_, x1), (t2_, x2) = fd[1e-6], fd[2e-6]
    out["richardson_h1e-6"] = {"dJdtheta": stats((4 * t1_ - t2_) / 3, ad_t), "dJdX": stats((4 * x1 - x2) / 3, ad_x)}
    trunc = (x2 - x1) / 3
    out["dJdX_h1e-6_error_vs_truncation_estimate"] = float(np.max(np.abs((x1 - ad_x) - trunc) / np.abs(trunc)))
    log(f"Richardson h = 1e-6: dJ/dX max rel {out['richardson_h1e-6']['dJdX']['max_entry_rel_err']:.2e}")
    if cpu is not None:
        cb = cpu(g=g, dh=dh, theta=theta, colors=colors, colptr=K.colptr, rowval=K.rowval,
                 prescribed=ch.prescribed_dofs, ydof=ydof, rtol=args.rtol)
        if cb:
            # the CPU restatement's J and gradients against the device's (both CG to the same rtol)
            cb["J_rel_diff_vs_device"] = abs(cb.pop("J") - J) / abs(J)
            cb["dJdtheta_rel_diff_vs_device"] = float(np.linalg.norm(cb.pop("dJdtheta") - dth_cg) / np.linalg.norm(dth_cg))
            cb["dJdX_rel_diff_vs_device"] = float(np.linalg.norm(cb.pop("dJdX") - dx_cg) / np.linalg.norm(dx_cg))
            cb.pop("u", None)
            cpu_s = cb["assembly_s"] + cb["apply_s"] + cb["solve_s"] + cb["J_and_gradients_s"]
            cb["value"] = cpu_s
            out["gpu_J_and_gradients_s"] = tj
            out["speedup_vs_cpu"] = cpu_s / tj if cb["solve_complete"] else None
        out["cpu_baseline"] = cb
    return out


def batch(args, cpu=None):
    import numpy as np
    import ferrite_amd as fa
    t0 = time.perf_counter()
    b, specs = fa.make_batch(args.problems, seed=args.seed)
    log(f"setup {time.perf_counter() - t0:.1f}s: {len(specs)} problems, {b.ncells} cells, {b.ndofs} dofs")
    shape = b.autotune("mixed")
    out = {"workload": f"{args.problems} batched Poisson problems (2D/3D Q1, log-uniform [2e3, 2e5] cells)",
           "cells": b.ncells, "dofs": b.ndofs, "autotuned_cells_per_wg_chunk": list(shape)}
    for prec in ("mixed", "fp64"):
        runs = [b.run(precision=prec, rtol=1e-8) for _ in range(args.reps)]
        r = min(runs, key=lambda x: x.total_ms)
        out[prec] = {"total_ms": r.total_ms, "assemble_ms": r.assemble_ms, "apply_ms": r.apply_ms,
                     "solve_ms": r.solve_ms, "cells_per_s": b.ncells / (r.total_ms * 1e-3),
                     "assembly_cells_per_s": b.ncells / (r.assemble_ms * 1e-3),
                     "p50_ms": r.percentile(50), "p99_ms": r.percentile(99), "converged": r.converged,
                     "total_iterations": r.total_iterations,
                     "all_total_ms": [x.total_ms for x in runs]}
        log(f"{prec}: total {r.total_ms:.2f} ms (assemble {r.assemble_ms:.2f}, solve {r.solve_ms:.2f}), "
            f"p50 {r.percentile(50):.2f} p99 {r.percentile(99):.2f} ms, converged {r.converged}")
    if cpu is not None:
        b.close()
        out["cpu_baseline"] = cpu(setups=specs)
        if out["cpu_baseline"]:
            out["speedup_vs_cpu"] = out["mixed"]["cells_per_s"] / out["cpu_baseline"]["value"]
    return out


def heat(args, cpu=None):
    import numpy as np
    import ferrite_amd as fa
    g = fa.generate_grid("Quadrilateral", (args.nx, args.ny))
    ip = fa.Lagrange("RefQuadrilateral", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefQuadrilateral", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    ch = fa.ConstraintHandler(dh)
    facets = [fc for name in ("left", "right", "bottom", "top") for fc in g.facetsets[name]]
    ch.add(fa.Dirichlet("u", facets, lambda x, t: 0.0)).close()
    el = fa.PoissonElement()
    kms = []

    def step():
        fa.assemble_global(fa.start_assemble(K, f), cv, colors, el)
        kms.append(fa.timings(K)["assemble_kernel_ms"])
        fa.apply(K, f, ch)
        return fa.solve(K, f, mode="cg", rtol=args.rtol)

    step()                      # solver set-up and graph capture
    reps = max(args.reps, 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        res = step()
    dt = (time.perf_counter() - t0) / reps
    ms = min(kms)
    out = {"workload": f"heat equation Q1 quad {args.nx}x{args.ny} on [-1,1]^2, 2x2 Gauss, u = 0 on the boundary",
           "cells": g.ncells, "dofs": dh.ndofs, "nnz": K.nnz, "colors": len(colors),
           "assembly_kernel_ms": ms, "assembly_kernel_ms_all": kms, "assembly_cells_per_s": g.ncells / (ms * 1e-3),
           "ms_per_step": dt * 1e3, "cells_per_s": g.ncells / dt,
           "step": "assemble + apply! + CG to rtol with u downloaded, after one untimed step",
           "cg_iterations": res.iterations, "relres": res.relres, "u_norm": float(np.linalg.norm(res.u))}
    log(f"heat {args.nx}x{args.ny}: assembly {ms * 1e3:.1f} us, step {dt * 1e3:.2f} ms "
        f"({res.iterations} its), |u| = {out['u_norm']:.12f}")
    if cpu is not None:
        out["cpu_baseline"] = cpu(g=g, dh=dh, colors=colors, colptr=K.colptr, rowval=K.rowval,
                                  prescribed=ch.prescribed_dofs)
        if out["cpu_baseline"]:
            out["assembly_speedup_vs_cpu"] = out["assembly_cells_per_s"] / out["cpu_baseline"]["value"]
            out["step_speedup_vs_cpu"] = out["cells_per_s"] / out["cpu_baseline"]["assemble_solve_cells_per_s"]
    return out


DEFAULTS = {"heat": dict(nx=100, ny=100, rtol=1e-10), "elasticity": dict(n=48, rtol=1e-8),
            "cantilever": dict(nx=512, ny=256, rtol=2e-9), "batch": dict(problems=64, rtol=1e-8)}


def default_args(config, **over):
    """The argument namespace of `config` with its defaults (and `over` on top)."""
    ns = argparse.Namespace(n=48, nx=512, ny=256, fd=64, problems=64, reps=3, seed=0, rtol=0.0)
    for k, v in {**DEFAULTS[config], **over}.items():
        setattr(ns, k, v)
    return ns


RUNS = {"heat": heat, "elasticity": elasticity, "cantilever": cantilever, "batch": batch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=tuple(RUNS))
    ap.add_argument("--n", type=int)
    ap.add_argument("--nx", type=int)
    ap.add_argument("--ny", type=int)
    ap.add_argument("--fd", type=int)
    ap.add_argument("--problems", type=int)
    ap.add_argument("--reps", type=int)
    ap.add_argument("--seed", type=int)
    ap.add_argument("--rtol", type=float,
                    help="CG relative residual (default 1e-10 heat, 1e-8 elasticity and batch, 2e-9 cantilever)")
    a = ap.parse_args()
    args = default_args(a.config, **{k: v for k, v in vars(a).items() if v is not None and k != "config"})
    print(json.dumps({"config": a.config, **RUNS[a.config](args)}), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
