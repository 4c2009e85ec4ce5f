"""Headline benchmark: 3D Poisson −∇·(k∇u) = 1 on a 128³ Q1 hex grid, assemble + apply! + solve.

One step = fused colored assembly of K (nzval) and f on the device, Dirichlet apply!
on 'left' ∪ 'right', and a CG solve to ‖f − K u‖/‖f‖ ≤ 1e-10, with all inputs resident
in HBM.  Two forms are timed back to back on the same FP64-assembled K:
  * fp64  : FP64 Jacobi-PCG on the solver's general sliced-ELL copy of K
  * mixed : FP32-SpMV Jacobi-PCG (FP32 vectors, FP64 dot products) with FP64 iterative
            refinement by reliable updates (true residual f − K u in FP64), on the cell form of
            the operator the structured mesh of congruent cells allows (per-cell k_e times one
            element matrix, used after an entry-by-entry check against K)
Timed too and reported beside: the mixed solver on the lattice value planes (`mixed_lattice`),
the FP64 solver on the planes (`fp64_lattice`) and on the cell form (`fp64_cells`), both solvers
on the element matrices of the assembly (`mixed_elements`, `fp64_elements`) and the mixed solver on
the sliced copy (`mixed_sell`), so `mixed_speedup_same_operator` compares the two solvers on each
operator form; and (`distorted`) the same workload on a mesh with moved interior nodes, where only
the sliced copy and the element matrices apply.
`value` is the mixed form's cells/s (the headline); the FP64 form's cells/s, their
ratio and ‖u_mixed − u_fp64‖/‖u_fp64‖ are reported beside it.  The FP32-element-math
assembly variant is timed separately with its K and u errors (an FP32 Ke moves u by
~1e-6 at this conditioning, above the 1e-8 bar, so it is not used for the headline).  Per-cell conductivity
k_e = exp(g_e), g_e ~ N(0, 0.5²) i.i.d. from the seed (synthetic data).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu-baseline]
       python bench.py --config heat|elasticity|cantilever|batch [--no-cpu-baseline]

`--config` runs one of the other workloads instead (ferrite_amd.profile_configs: the 100x100 quad
heat equation, Q2 hex elasticity at 48³, the 512x256 cantilever compliance with its gradients,
the batch of 64 Poisson problems) and adds the timed CPU restatement of the same workload as
`cpu_baseline`; it prints that workload's JSON record, not the headline line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "ferrite.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the roofline kernel at this workload, from rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE runs calibrated against a 1 GiB device copy (see the profile summary it names)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_q1hex_128.json")
ASSEMBLY_BYTES_PER_CELL = 1508.6   # Q1 hex FP64 colored scatter, algorithmic


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--coloring", default="workstream")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--config", choices=("heat", "elasticity", "cantilever", "batch"), default=None,
                    help="run another workload with its CPU baseline instead of the headline")
    return ap.parse_args()


def measured_traffic(kernel, n):
    """Calibrated HBM bytes per launch of `kernel` at n^3, if a committed PMC summary holds them."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    e = t.get(kernel)
    if not e or e.get("n") != n:
        return None
    return e.get("bytes_per_launch")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU (a multi-GPU run would partition the "
                                   "cells into slabs with a one-plane halo; not built); run with --gpus 1"}))
        sys.exit(2)
    if args.config:
        run_config(args)
        return
    import numpy as np
    import torch

    import ferrite_amd as fa

    n = args.n
    nel = (n, n, n)
    t0 = time.time()
    g = fa.generate_grid("Hexahedron", nel, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    t_host = time.perf_counter()
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh, device=0)
    t_host = time.perf_counter() - t_host
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g, args.coloring)
    rng = np.random.default_rng(args.seed)
    k = np.exp(rng.normal(0.0, 0.5, size=g.ncells))
    el = fa.PoissonElement(k=k, source=1.0)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["right"], lambda x, t: 0.0)).close()
    log(f"setup {time.time() - t0:.1f}s: {g.ncells} cells, {dh.ndofs} dofs, nnz {K.nnz}, {len(colors)} colors")
    # close!(dh) + allocate_matrix(dh) on the device (fem_dofs_distribute, fem_pattern_build),
    # copied out for the host mirror, against the host restatement just run
    torch.cuda.synchronize()
    t_dev = time.perf_counter()
    dh_d = fa.DofHandler(g).add("u", ip).close(device=0)
    K_d = fa.allocate_matrix(fa.MI355XCSC, dh_d)
    t_dev = time.perf_counter() - t_dev
    same = bool(np.array_equal(dh_d.cell_dofs, dh.cell_dofs) and np.array_equal(K_d.colptr, K.colptr)
                and np.array_equal(K_d.rowval, K.rowval))
    setup = {"host_s": t_host, "device_s": t_dev, "device_equals_host": same,
             "what": "close!(dh) + allocate_matrix(dh) pattern + device upload; the device path also "
                     "copies cell_dofs, colptr and rowval back for the host mirror"}
    log(f"setup: host {t_host:.3f} s, device {t_dev:.3f} s, identical {same}")
    del K_d, dh_d

    def step(mode, asm_precision="fp64"):
        A = fa.start_assemble(K, f)
        fa.assemble_global(A, cv, colors, el, precision=asm_precision)
        t_asm = fa.timings(K)
        fa.apply(K, f, ch)
        res = fa.solve(K, f, mode=mode, rtol=args.rtol, download=False)
        return res, t_asm, fa.timings(K)

    results = {}
    for name, mode in (("fp64", "cg"), ("mixed", "mixed")):
        for _ in range(args.warmup):
            step(mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        iters = []
        asm_ms = []
        for _ in range(args.steps):
            res, ta, ts = step(mode)
            iters.append((res.iterations, res.outer_iterations, res.relres))
            asm_ms.append(ta["assemble_kernel_ms"])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / args.steps
        u = fa.device.DeviceVector(K, 2).to_numpy()
        # one more step with the CG kernels launched directly and every SpMV bracketed by HIP
        # events on the context stream: the SpMV's launch duration inside the solve (same
        # kernels and operands as the timed graphs), used for the roofline below
        fa.set_tuning(K, time_kernels=1)
        _, _, ts = step(mode)
        fa.set_tuning(K, time_kernels=0)
        spmv_ms, spmv_n = ts["spmv_ms"], ts["spmv_launches"]
        results[name] = dict(ms=dt * 1e3, cells_per_s=g.ncells / dt, iters=iters, u=u, operator=res.operator,
                             unit_diagonal=res.unit_diagonal, cell_shape=res.cell_shape,
                             asm_kernel_ms=float(np.median(asm_ms)), spmv_avg_ms=spmv_ms / max(spmv_n, 1),
                             spmv_launches=spmv_n)
        log(f"{name}: {dt * 1e3:.1f} ms/step, iters {iters[-1]}, assembly kernels {np.median(asm_ms):.3f} ms, "
            f"SpMV {spmv_ms / max(spmv_n, 1) * 1e3:.1f} us avg over {spmv_n} launches")

    uerr = float(np.linalg.norm(results["mixed"]["u"] - results["fp64"]["u"]) / np.linalg.norm(results["fp64"]["u"]))

    def context(name, mode, **tuning):
        """Context, not the headline: one solver on another operator form (same K, same steps),
        with its SpMV's in-solve launch duration from one more step with timed SpMVs."""
        fa.set_tuning(K, **tuning)
        for _ in range(args.warmup):
            step(mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r, _, _ = step(mode)
        torch.cuda.synchronize()
        dtc = (time.perf_counter() - t1) / args.steps
        fa.set_tuning(K, time_kernels=1)
        _, _, ts = step(mode)
        fa.set_tuning(K, lattice=0, cell_operator=0, elem_operator=0, time_kernels=0)
        log(f"{name}: {dtc * 1e3:.1f} ms/step, {r.iterations} iterations, operator {r.operator}, SpMV "
            f"{ts['spmv_ms'] / max(ts['spmv_launches'], 1) * 1e3:.1f} us")
        return {"cells_per_s": g.ncells / dtc, "ms_per_step": dtc * 1e3, "operator": r.operator,
                "cg_iterations": r.iterations, "relres": r.relres,
                "spmv_avg_ms": ts["spmv_ms"] / max(ts["spmv_launches"], 1), "unit_diagonal": r.unit_diagonal}

    mixed_lattice = context("mixed on the lattice planes", "mixed", cell_operator=-1)
    fp64_lattice = context("fp64 on the lattice planes", "cg", lattice=1, cell_operator=-1)
    fp64_cells = context("fp64 on the cell form", "cg", cell_operator=1)
    # like for like on the general operator forms, which any mesh gets: the sliced ELL copy of K
    # and the element matrices of the owner-computes assembly
    mixed_sell = context("mixed on the sliced copy", "mixed", lattice=-1, cell_operator=-1, elem_operator=-1)
    mixed_elements = context("mixed on the element matrices", "mixed", lattice=-1, cell_operator=-1, elem_operator=1)
    fp64_elements = context("fp64 on the element matrices", "cg", lattice=-1, cell_operator=-1, elem_operator=1)
    fa.set_tuning(K, lattice=0, cell_operator=0, elem_operator=0)

    # FP32 element math with FP64 accumulation: K error and the u it leads to (reported beside)
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="fp64")
    nz64 = K.nzval()
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="mixed")
    asm32_ms = fa.timings(K)["assemble_kernel_ms"]
    nz32 = K.nzval()
    kerr32 = float(np.linalg.norm(nz32 - nz64) / np.linalg.norm(nz64))
    fa.apply(K, f, ch)
    r32 = fa.solve(K, f, mode="mixed", rtol=args.rtol)
    uerr32 = float(np.linalg.norm(r32.u - results["fp64"]["u"]) / np.linalg.norm(results["fp64"]["u"]))
    # leave K in its FP64-assembled, constrained state for the SpMV timing below
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="fp64")
    fa.apply(K, f, ch)

    # roofline of the dominant kernel: the FP32 SpMV of the mixed solver, algorithmic bytes of its
    # operator form.  Cell form (z-marching): t = S and x read over the padded lattice, k_e over
    # it (one per cell origin row), y written over the marched planes.  Lattice planes: 14 FP32
    # value planes over the computed rows (13 with the unit diagonal, whose centre plane is not
    # read), x over the padded lattice, y over the computed rows; backward neighbours re-read
    # planes another row already read.  Sliced ELL: y = K x with int32 indices.
    spmv_iso_ms = fa.spmv_bench(K, mixed=True, reps=50)   # back to back, operator partly cache-resident
    nnz, nd = K.nnz, dh.ndofs
    px = n + 3                                               # n + 1 nodes per axis, one padding node each side
    npad, omax = px ** 3, 1 + px + px * px

    def lattice_roofline(unit, ms):
        kern = f"k_lat_spmv<float, float, 14, 1, {'true' if unit else 'false'}>"
        nbytes = (npad - 2 * omax) * ((13 if unit else 14) * 4 + 4) + 4 * npad
        gbs = nbytes / (ms * 1e-3) / 1e9
        return kern, nbytes, gbs

    spmv_ms = results["mixed"]["spmv_avg_ms"]               # inside the CG loop of the timed steps
    if results["mixed"]["operator"] == "cells" and results["mixed"]["cell_shape"]:
        # the z-marching template the solver reports (MAP, ZN, KN, stages, rows per thread)
        SPMV32_KERNEL = "k_cell_spmv_z<float, {}, {}, {}, {}, {}>".format(*results["mixed"]["cell_shape"])
        spmv_bytes = 3 * 4 * npad + 4 * (px - 2) * px * px
    elif results["mixed"]["operator"] == "cells":
        SPMV32_KERNEL = "k_cell_spmv<float, 3, 1, 4>"   # tile form (lattice too wide for the march)
        spmv_bytes = 3 * 4 * npad + 4 * (px - 2) * px * px
    elif results["mixed"]["operator"] == "lattice":
        SPMV32_KERNEL, spmv_bytes, _ = lattice_roofline(results["mixed"]["unit_diagonal"], spmv_ms)
    else:
        SPMV32_KERNEL = "k_sell<float, float, 0, true, 8, false, true>"   # sliced ELL, u16 offsets, NT loads
        spmv_bytes = nnz * (4 + 4) + 4 * (nd + 1) + 4 * nd + 4 * nd
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    stream = same_size_stream(torch, npad)
    lat_kernel, lat_bytes, lat_gbs = lattice_roofline(mixed_lattice["unit_diagonal"], mixed_lattice["spmv_avg_ms"])
    spmv64_iso_ms = fa.spmv_bench(K, mixed=False, reps=50)
    spmv64_ms = results["fp64"]["spmv_avg_ms"]
    spmv64_bytes = nnz * (8 + 4) + 4 * (nd + 1) + 8 * nd + 8 * nd
    asm_ms = results["fp64"]["asm_kernel_ms"]
    asm_gbs = ASSEMBLY_BYTES_PER_CELL * g.ncells / (asm_ms * 1e-3) / 1e9

    distorted = distorted_run(args, fa, np, torch, g, ip, cv, dh, colors, k)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, g, dh, k, colors, K, ch)

    out = {
        "metric": "cells/s assemble+solve Q1 hex 128³: FP64 plain vs mixed @ u rel err ≤1e-8",
        "value": results["mixed"]["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results["mixed"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 assembly; fp32 SpMV/vectors + fp64 refinement and dot products",
        "data": "synthetic (seeded k_e = exp(N(0, 0.5^2)))",
        "config": {"workload": f"3D Poisson Q1 hex {n}^3 on [0,1]^3, 2x2x2 Gauss, Dirichlet left+right, "
                               f"CG to rel. residual {args.rtol:g}",
                   "cells": g.ncells, "dofs": dh.ndofs, "nnz": nnz, "colors": len(colors),
                   "coloring": args.coloring, "seed": args.seed},
        "fp64": {"cells_per_s": results["fp64"]["cells_per_s"], "ms_per_step": results["fp64"]["ms"],
                 "operator": results["fp64"]["operator"],
                 "cg_iterations": results["fp64"]["iters"][-1][0], "relres": results["fp64"]["iters"][-1][2]},
        "mixed": {"cells_per_s": results["mixed"]["cells_per_s"], "ms_per_step": results["mixed"]["ms"],
                  "operator": results["mixed"]["operator"],
                  "cg_iterations": results["mixed"]["iters"][-1][0],
                  "refinements": results["mixed"]["iters"][-1][1], "relres": results["mixed"]["iters"][-1][2]},
        "mixed_speedup_vs_fp64": results["fp64"]["ms"] / results["mixed"]["ms"],
        "mixed_lattice": mixed_lattice,
        "fp64_lattice": fp64_lattice,
        "fp64_cells": fp64_cells,
        "mixed_speedup_vs_fp64_lattice": fp64_lattice["ms_per_step"] / results["mixed"]["ms"],
        "mixed_speedup_vs_fp64_cells": fp64_cells["ms_per_step"] / results["mixed"]["ms"],
        "mixed_sell": mixed_sell,
        "mixed_elements": mixed_elements,
        "fp64_elements": fp64_elements,
        # mixed vs FP64 on the same operator form (FP64 on the sliced copy = the "fp64" line)
        "mixed_speedup_same_operator": {
            "cells": fp64_cells["ms_per_step"] / results["mixed"]["ms"],
            "lattice": fp64_lattice["ms_per_step"] / mixed_lattice["ms_per_step"],
            "sell": results["fp64"]["ms"] / mixed_sell["ms_per_step"],
            "elements": fp64_elements["ms_per_step"] / mixed_elements["ms_per_step"]},
        "distorted": distorted,
        "setup": setup,
        "u_rel_err_mixed_vs_fp64": uerr,
        "fp32_element_assembly": {"kernel_ms": asm32_ms, "K_rel_err_vs_fp64": kerr32,
                                  "u_rel_err_vs_fp64": uerr32},
        "dofs_per_s": dh.ndofs / (results["mixed"]["ms"] * 1e-3),
        "roofline": {"bound": "hbm", "kernel": SPMV32_KERNEL,
                     "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": spmv_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(SPMV32_KERNEL, n),
                     "bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_ms,
                     "launches": results["mixed"]["spmv_launches"], "isolated_avg_launch_ms": spmv_iso_ms,
                     "same_size_stream": stream},
        # the same workload's SpMV on the lattice value planes (the mixed_lattice run)
        "roofline_lattice": {"bound": "hbm", "kernel": lat_kernel, "achieved": lat_gbs, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": lat_gbs / HBM_PEAK_GBS,
                             "traffic": measured_traffic(lat_kernel, n), "bytes_per_launch": lat_bytes,
                             "avg_launch_ms": mixed_lattice["spmv_avg_ms"]},
        "spmv_fp64": {"avg_launch_ms": spmv64_ms, "isolated_avg_launch_ms": spmv64_iso_ms,
                      "bytes_per_launch": spmv64_bytes,
                      "achieved_GBs": spmv64_bytes / (spmv64_ms * 1e-3) / 1e9},
        "assembly_fp64": {"kernel_ms": asm_ms, "bytes_per_cell": ASSEMBLY_BYTES_PER_CELL,
                          "achieved_GBs": asm_gbs, "frac": asm_gbs / HBM_PEAK_GBS,
                          "cells_per_s": g.ncells / (asm_ms * 1e-3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def distorted_run(args, fa, np, torch, g0, ip, cv, dh0, colors, k):
    """The same workload on a distorted mesh: interior nodes moved by up to 0.3 h (seeded), so
    cells are general trilinear hexahedra and the cell form does not apply; the lattice planes
    are switched off as well, so the solvers run on the operators any (unstructured) mesh gets:
    the sliced ELL copy of K and the element matrices.  Assembly takes the trilinear geometry
    form.  Returns per-solver cells/s, iterations and in-solve SpMV durations."""
    n = args.n
    g = fa.generate_grid("Hexahedron", (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    X = g.nodes
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += np.random.default_rng(args.seed + 1).uniform(-0.3 / n, 0.3 / n, (int(inner.sum()), 3))
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh, device=0)
    f = fa.allocate_vector(K)
    el = fa.PoissonElement(k=k, source=1.0)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["right"], lambda x, t: 0.0)).close()
    out = {"workload": f"as above on a distorted {n}^3 mesh (interior nodes moved up to 0.3 h), "
                       "lattice forms off", "seed": args.seed + 1}

    def step(mode):
        fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="fp64")
        t_asm = fa.timings(K)["assemble_kernel_ms"]
        fa.apply(K, f, ch)
        return fa.solve(K, f, mode=mode, rtol=args.rtol, download=False), t_asm

    us = {}
    for name, mode, elem in (("fp64_sell", "cg", -1), ("mixed_sell", "mixed", -1),
                             ("fp64_elements", "cg", 1), ("mixed_elements", "mixed", 1)):
        fa.set_tuning(K, lattice=-1, cell_operator=-1, elem_operator=elem)
        for _ in range(args.warmup):
            step(mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r, t_asm = step(mode)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / args.steps
        us[name] = fa.device.DeviceVector(K, 2).to_numpy()
        fa.set_tuning(K, lattice=-1, cell_operator=-1, elem_operator=elem, time_kernels=1)
        _ = step(mode)
        ts = fa.timings(K)
        spmv = ts["spmv_ms"] / max(ts["spmv_launches"], 1)
        out[name] = {"cells_per_s": g.ncells / dt, "ms_per_step": dt * 1e3, "operator": r.operator,
                     "cg_iterations": r.iterations, "relres": r.relres, "spmv_avg_ms": spmv,
                     "assembly_kernel_ms": t_asm}
        log(f"distorted {name}: {dt * 1e3:.1f} ms/step, {r.iterations} iterations, operator {r.operator}, "
            f"SpMV {spmv * 1e3:.1f} us, assembly {t_asm:.3f} ms")
    ref = us["fp64_sell"]
    for name in ("mixed_sell", "fp64_elements", "mixed_elements"):
        out[name]["u_rel_err_vs_fp64_sell"] = float(np.linalg.norm(us[name] - ref) / np.linalg.norm(ref))
    out["mixed_speedup_same_operator"] = {
        "sell": out["fp64_sell"]["ms_per_step"] / out["mixed_sell"]["ms_per_step"],
        "elements": out["fp64_elements"]["ms_per_step"] / out["mixed_elements"]["ms_per_step"]}
    # FP32 sliced-ELL SpMV roofline on this mesh (the same CSR byte count as on the box)
    nnz, nd = K.nnz, dh.ndofs
    b32 = nnz * (4 + 4) + 4 * (nd + 1) + 4 * nd + 4 * nd
    ms32 = out["mixed_sell"]["spmv_avg_ms"]
    out["roofline_sell32"] = {"bound": "hbm", "kernel": "k_sell<float, float, 0, true, 8, false, true>",
                              "achieved": b32 / (ms32 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": b32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": b32,
                              "avg_launch_ms": ms32}
    del K, f
    return out


def same_size_stream(torch, npad, reps=50):
    """What a plain streaming kernel of the cell-form SpMV's size achieves on this GPU: torch's
    addcmul y = a + b·c over FP32 vectors of npad entries (three read, one written: the SpMV's
    t, x, k_e in and y out).  Timed with events over `reps` launches, once on one set of
    vectors (back to back: they stay in the 256 MB Infinity Cache, like the isolated SpMV
    timing) and once rotating over eight sets (288 MB, more than that cache holds)."""
    sets = [[torch.rand(npad, device="cuda", dtype=torch.float32) for _ in range(4)] for _ in range(8)]
    nbytes = 16 * npad
    out = {"kernel": "torch.addcmul(a, b, c, out=y), FP32", "bytes_per_launch": nbytes}
    for name, pick in (("same_vectors", lambda i: sets[0]), ("rotating_8_sets", lambda i: sets[i % 8])):
        for i in range(8):
            a, b, c, y = pick(i)
            torch.addcmul(a, b, c, out=y)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for i in range(reps):
            a, b, c, y = pick(i)
            torch.addcmul(a, b, c, out=y)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"avg_launch_ms": ms, "GBs": nbytes / (ms * 1e-3) / 1e9}
    del sets
    torch.cuda.empty_cache()
    return out


def cpu_baseline(args, g, dh, k, colors, K, ch):
    """The oracle's C restatement of Ferrite's threaded colored assembly + apply! + PCG, host cores."""
    try:
        from oracle import cbaseline
        return cbaseline.run(g, dh, k, colors, K.colptr, K.rowval, ch.prescribed_dofs,
                             threads=args.cpu_threads, rtol=args.rtol, log=log)
    except Exception as e:  # the oracle's C library is built by __graft_entry__.build()
        log(f"cpu baseline unavailable: {e}")
        return None


def config_cpu_leg(args):
    """The CPU baseline of a `--config` workload: the oracle's C restatement of the same workload
    on the host cores, called by the workload after its device measurements."""
    if args.no_cpu_baseline:
        return None
    from oracle import cbaseline as CB
    nt = args.cpu_threads

    def leg(**kw):
        try:
            if args.config == "heat":
                return CB.run_heat(kw["g"], kw["dh"], kw["colors"], kw["colptr"], kw["rowval"], kw["prescribed"],
                                   threads=nt, log=log)
            if args.config == "elasticity":
                return CB.run_elasticity(kw["g"], kw["dh"], kw["E"], kw["nu"], kw["colors"], kw["colptr"],
                                         kw["rowval"], threads=nt, log=log)
            if args.config == "cantilever":
                return CB.run_cantilever(kw["g"], kw["dh"], kw["theta"], kw["colors"], kw["colptr"], kw["rowval"],
                                         kw["prescribed"], kw["ydof"], threads=nt, rtol=kw["rtol"], log=log)
            import ferrite_amd as fa
            from ferrite_amd.batch import poisson_problem
            probs = []
            for celltype, nel, k in kw["setups"]:
                g, dh, colors, ch, k = poisson_problem(celltype, nel, k)
                cp, rv = fa.create_sparsity_pattern(dh)
                probs.append((g, dh, colors, cp, rv, ch.prescribed_dofs, k))
            return CB.run_batch(probs, threads=nt, log=log)
        except Exception as e:  # the oracle's C library is built by __graft_entry__.build()
            log(f"cpu baseline unavailable: {e}")
            return None
    return leg


def run_config(args):
    from ferrite_amd import profile_configs as PC
    pc_args = PC.default_args(args.config)
    out = PC.RUNS[args.config](pc_args, cpu=config_cpu_leg(args))
    print(json.dumps({"config": args.config, **out}), flush=True)


if __name__ == "__main__":
    main()
