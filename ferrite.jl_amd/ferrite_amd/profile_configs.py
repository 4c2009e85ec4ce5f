"""Full-size runs of the elasticity, differentiation and batching workloads (one JSON line each).

    python -m ferrite_amd.profile_configs elasticity   [--n 48]
    python -m ferrite_amd.profile_configs cantilever   [--nx 512 --ny 256 --fd 64]
    python -m ferrite_amd.profile_configs batch        [--problems 64]

elasticity  Q2 hex^3 small-strain elasticity on generate_grid(Hexahedron, (n, n, n)) over [0,1]^3,
            3x3x3 Gauss, nu = 0.3, E_e ~ U[0.5, 2.0] (seed 0), clamped 'left', traction (0,0,-1)
            on 'right' (3x3 facet rule).  Times the FP64 element path and the FP32 MFMA element
            path (both with FP64 scatter), reports Ke / u differences between them and the
            element contraction rate (dense B^T D B count, 2,283,228 flop/cell, and the six
            27x27x27 products the kernel performs).
cantilever  Plane-strain Q1^2 cantilever on (nx, ny) cells over [0,2]x[0,1], E = E_min +
            theta^3 (1 - E_min), theta ~ U[0.2, 1.0] (seed 0), clamped at x = 0, point load
            (0, -1) at (2, 0.5).  Times J and the adjoint gradients; checks `fd` random entries
            of dJ/dtheta and of dJ/dX against central differences (h = 1e-6 and 1e-4; FP64 CG
            attains a relative residual of about 1e-9 on this system, which bounds how well a
            difference of two J values resolves the small entries).
batch       `problems` seeded scalar Poisson problems (2D/3D, log-uniform cell counts in
            [2e3, 2e5]) assembled and solved (mixed CG to 1e-8) in one batched context after
            the launch-shape autotuner; total cells/s and p50 / p99 completion times.
"""
import argparse
import json
import os
import sys
import time

DENSE_BTDB_FLOP_PER_CELL = 2_283_228          # 27 q x 2 x (6*6*81 + 81*6*81)
PRODUCT_FLOP_PER_CELL = 6 * 27 ** 3 * 2       # six 27x27 products over 27 points
MFMA_FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def elasticity(args):
    import numpy as np
    import ferrite_amd as fa
    n = args.n
    t0 = time.perf_counter()
    g = fa.generate_grid("Hexahedron", (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ip = fa.Lagrange("RefHexahedron", 2) ** 3
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 3), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    E = np.random.default_rng(args.seed).uniform(0.5, 2.0, g.ncells)
    el = fa.ElasticityElement(E=E, nu=0.3)
    fv = fa.FacetValues(fa.FacetQuadratureRule("RefHexahedron", 3), ip)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"], lambda x, t: [0.0, 0.0, 0.0])).close()
    log(f"setup {time.perf_counter() - t0:.1f}s: {g.ncells} cells, {dh.ndofs} dofs, nnz {K.nnz}, "
        f"{len(colors)} colors")
    out = {"workload": f"Q2 hex^3 elasticity {n}^3, 3x3x3 Gauss, E ~ U[0.5,2], nu 0.3",
           "cells": g.ncells, "dofs": dh.ndofs, "nnz": K.nnz, "colors": len(colors)}
    sols, nzs = {}, {}
    for prec in ("fp64", "mixed"):
        ms = []
        for _ in range(args.reps):
            fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
            ms.append(fa.timings(K)["assemble_kernel_ms"])
        best = min(ms)
        nzs[prec] = K.nzval()
        fa.apply_traction(K, f, fv, g.facetsets["right"], (0.0, 0.0, -1.0))
        fa.apply(K, f, ch)
        t1 = time.perf_counter()
        res = fa.solve(K, f, mode="cg", rtol=args.rtol)
        ts = time.perf_counter() - t1
        sols[prec] = res.u
        out[prec] = {"assembly_ms": best, "assembly_ms_all": ms, "cells_per_s": g.ncells / (best * 1e-3),
                     "dense_btdb_tflops": DENSE_BTDB_FLOP_PER_CELL * g.ncells / (best * 1e-3) / 1e12,
                     "product_tflops": PRODUCT_FLOP_PER_CELL * g.ncells / (best * 1e-3) / 1e12,
                     "solve_s": ts, "cg_iterations": res.iterations, "relres": res.relres}
        log(f"{prec}: assembly {best:.3f} ms, CG {res.iterations} its {ts:.2f}s relres {res.relres:.2e}")
    a, b = nzs["mixed"], nzs["fp64"]
    out["K_rel_err_mixed_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["u_rel_err_mixed_vs_fp64"] = float(np.linalg.norm(sols["mixed"] - sols["fp64"]) /
                                           np.linalg.norm(sols["fp64"]))
    out["mfma_fp32_frac_product_count"] = out["mixed"]["product_tflops"] / MFMA_FP32_PEAK_TFLOPS
    out["fp64_frac_product_count"] = out["fp64"]["product_tflops"] / FP64_PEAK_TFLOPS
    return out


def cantilever(args):
    import numpy as np
    import ferrite_amd as fa
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
    X0 = g.nodes.copy()
    t1 = time.perf_counter()
    J, dth, dx, res = prob.compliance(theta)
    tj = time.perf_counter() - t1
    t1 = time.perf_counter()
    fa.compliance_gradient(K, theta)
    tg = time.perf_counter() - t1
    log(f"J = {J:.10e}: forward {tj:.2f}s ({res.iterations} CG its), adjoint kernel {tg * 1e3:.2f} ms")
    pick = np.random.default_rng(args.seed + 1)
    cells = pick.choice(g.ncells, args.fd, replace=False)
    coords = pick.choice(g.nnodes * 2, args.fd, replace=False)
    fd_rows = {}
    for h in (1e-6, 1e-4):
        rows_t, rows_x = [], []
        for e in cells:
            tp, tm = theta.copy(), theta.copy()
            tp[e] += h
            tm[e] -= h
            fd = (prob.compliance(tp)[0] - prob.compliance(tm)[0]) / (2 * h)
            rows_t.append((fd, dth[e]))
        for k in coords:
            nd, d = divmod(int(k), 2)
            Xp, Xm = X0.copy(), X0.copy()
            Xp[nd, d] += h
            Xm[nd, d] -= h
            fd = (prob.compliance(theta, Xp)[0] - prob.compliance(theta, Xm)[0]) / (2 * h)
            rows_x.append((fd, dx[nd, d]))
        fd_rows[h] = (rows_t, rows_x)
        log(f"h = {h:g}: {len(rows_t)} + {len(rows_x)} central differences done")
    prob.compliance(theta, X0)

    def stats(rows):
        fd = np.array([r[0] for r in rows])
        ad = np.array([r[1] for r in rows])
        rel = np.abs(fd - ad) / np.maximum(np.abs(ad), 1e-300)
        return {"entries": len(rows), "normwise_rel_err": float(np.linalg.norm(fd - ad) / np.linalg.norm(ad)),
                "max_entry_rel_err": float(rel.max()), "median_entry_rel_err": float(np.median(rel)),
                "entries_above_1e-6": int(np.sum(rel > 1e-6)),
                "entries_above_1e-6_and_floor": int(np.sum(np.abs(fd - ad) > 1e-6 * np.abs(ad) + 1e-7 * abs(J)))}
    out = {"workload": f"plane-strain Q1^2 cantilever {nx}x{ny}, SIMP p=3, E_min 1e-3, point load at node {tip}",
           "cells": g.ncells, "dofs": dh.ndofs, "J": J, "forward_s": tj, "cg_iterations": res.iterations,
           "cg_rtol": args.rtol, "relres": res.relres, "adjoint_kernel_ms": tg * 1e3}
    for h, (rows_t, rows_x) in fd_rows.items():
        out[f"fd_h{h:g}"] = {"dJdtheta": stats(rows_t), "dJdX": stats(rows_x)}
    return out


def batch(args):
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
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=("elasticity", "cantilever", "batch"))
    ap.add_argument("--n", type=int, default=48)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--fd", type=int, default=64)
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rtol", type=float, default=0.0,
                    help="CG relative residual (default 1e-8 elasticity, 2e-9 cantilever)")
    args = ap.parse_args()
    fn = {"elasticity": elasticity, "cantilever": cantilever, "batch": batch}[args.config]
    if args.rtol == 0.0:
        args.rtol = {"elasticity": 1e-8, "cantilever": 2e-9, "batch": 1e-8}[args.config]
    print(json.dumps({"config": args.config, **fn(args)}), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
