"""Kernel-level profiling driver for rocprofv3 (counters and per-kernel durations).

    python -m ferrite_amd.profile_kernels [--n 128] [--reps 5]

Builds the 3D Poisson Q1 hex n³ system of the headline benchmark, then runs, in order:
  1. a plain device copy of a known size (torch), the calibration for FETCH_SIZE/WRITE_SIZE;
  2. `reps` fused assemblies (FP64 element kernel), apply!;
  3. `reps` FP64 and FP32 solver SpMVs.
Each phase prints a line, so a counter run can be matched to its kernels by name.
"""
import argparse
import os
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--sweep", action="store_true", help="time SpMV / assembly launch shapes")
    ap.add_argument("--solve", action="store_true", help="also run one FP64 and one mixed solve")
    ap.add_argument("--cg-sweep", action="store_true",
                    help="CG vector step: elements per thread (solve loop times, both solvers)")
    ap.add_argument("--gather-sweep", action="store_true",
                    help="owner-computes assembly: cells loaded per column lane before adding")
    ap.add_argument("--ir-sweep", action="store_true",
                    help="mixed solver: reliable-update factor δ (inner_rtol) vs iterations and loop time")
    ap.add_argument("--cell-sweep", action="store_true", help="cell-form SpMV: XCD map x block counts")
    ap.add_argument("--lattice-sweep", action="store_true",
                    help="lattice vs sliced-ELL solver operator x block mapping: isolated and in-solve SpMV")
    ap.add_argument("--sell-sweep", action="store_true",
                    help="sliced-ELL unroll depth x slice mapping: isolated and in-solve SpMV durations")
    args = ap.parse_args()
    import numpy as np
    import torch

    import ferrite_amd as fa

    n = args.n
    g = fa.generate_grid("Hexahedron", (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ip = fa.Lagrange("RefHexahedron", 1)
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 2), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    k = np.exp(np.random.default_rng(0).normal(0.0, 0.5, size=g.ncells))
    el = fa.PoissonElement(k=k)
    ch = fa.ConstraintHandler(dh)
    ch.add(fa.Dirichlet("u", g.facetsets["left"] + g.facetsets["right"], lambda x, t: 0.0)).close()

    nbytes = args.copy_mib << 20
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"phase copy: {nbytes} bytes read + {nbytes} bytes written", flush=True)
    del a, b

    for _ in range(args.reps):
        fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="fp64")
    t = fa.timings(K)
    print(f"phase assemble: {args.reps} x {g.ncells} cells, last {t['assemble_kernel_ms']:.3f} ms "
          f"over {t['assemble_launches']} colour launches", flush=True)
    fa.apply(K, f, ch)
    ms64 = fa.spmv_bench(K, mixed=False, reps=args.reps)
    ms32 = fa.spmv_bench(K, mixed=True, reps=args.reps)
    print(f"phase spmv: fp64 {ms64:.4f} ms, fp32 {ms32:.4f} ms per launch (nnz {K.nnz}, n {dh.ndofs})",
          flush=True)
    if args.solve:
        # default operators (FP64: sliced ELL; mixed: cell form), then the mixed solver on the
        # lattice planes and the FP64 solver on the cell form
        for mode, tune in (("cg", {}), ("mixed", {}), ("mixed", {"cell_operator": -1}), ("cg", {"cell_operator": 1})):
            fa.set_tuning(K, **tune)
            r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
            print(f"phase solve {mode} [{r.operator}]: {r.iterations} iterations, relres {r.relres:.2e}", flush=True)
        fa.set_tuning(K, cell_operator=0)
    if args.cg_sweep:
        for split, ce in ((-1, 8), (0, 4), (0, 8), (0, 16)):
            fa.set_tuning(K, cg_elems=ce, cg_split=split)
            line = []
            for mode in ("cg", "mixed"):
                best = 1e30
                for _ in range(2):
                    r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
                    best = min(best, fa.timings(K)["solve_ms"])
                line.append(f"{mode} loop {best:.2f} ms ({r.iterations} its, {best / r.iterations * 1e3:.1f} us/it)")
            print(f"cg split {split} elems {ce}: {'; '.join(line)}", flush=True)
        fa.set_tuning(K, cg_elems=0, cg_split=0)
    if args.gather_sweep:
        for chunks in (1, 2, 4, 8, 16, 32):
            fa.set_tuning(K, gather_chunks=chunks)
            ts = []
            for prec in ("fp64", "mixed"):
                best = 1e30
                for _ in range(3):
                    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
                    best = min(best, fa.timings(K)["assemble_kernel_ms"])
                ts.append(best)
            print(f"gather chunks {chunks}: fp64 {ts[0]:.3f} ms, fp32-element {ts[1]:.3f} ms", flush=True)
        fa.set_tuning(K, gather_chunks=0)
        for grp in (1, 2, 4, 8):
            fa.set_tuning(K, gather_group=grp)
            ts = []
            for prec in ("fp64", "mixed"):
                best = 1e30
                for _ in range(3):
                    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
                    best = min(best, fa.timings(K)["assemble_kernel_ms"])
                ts.append(best)
            print(f"gather group {grp}: fp64 {ts[0]:.3f} ms, fp32-element {ts[1]:.3f} ms", flush=True)
        fa.set_tuning(K, gather_group=0)
        fa.apply(K, f, ch)
    if args.ir_sweep:
        for d in (0.3, 0.1, 0.03, 0.01, 0.003):
            best, r = 1e30, None
            for _ in range(2):
                r = fa.solve(K, f, mode="mixed", rtol=1e-10, inner_rtol=d, download=False)
                best = min(best, fa.timings(K)["solve_ms"])
            print(f"ir delta {d:g}: {r.iterations} its, {r.outer_iterations} reliable updates, loop {best:.2f} ms, "
                  f"relres {r.relres:.2e}", flush=True)
    if args.lattice_sweep:
        for lat, lname in ((-1, "sliced ELL"), (1, "lattice")):
            for smap, mname in ((1, "grid-stride"), (2, "XCD ranges")):
                fa.set_tuning(K, lattice=lat, sell_map=smap)
                a64 = fa.spmv_bench(K, mixed=False, reps=20)
                a32 = fa.spmv_bench(K, mixed=True, reps=20)
                fa.set_tuning(K, lattice=lat, sell_map=smap, time_kernels=1)
                line = []
                for mode in ("cg", "mixed"):
                    r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
                    t = fa.timings(K)
                    line.append(f"{mode} [{r.operator}] {t['spmv_ms'] / max(t['spmv_launches'], 1) * 1e3:.1f} us x "
                                f"{t['spmv_launches']} ({r.iterations} its, loop {t['solve_ms']:.1f} ms)")
                fa.set_tuning(K, lattice=lat, sell_map=smap, time_kernels=0)
                loops = []
                for mode in ("cg", "mixed"):
                    r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
                    loops.append(f"{mode} {fa.timings(K)['solve_ms']:.1f} ms")
                print(f"operator {lname} {mname}: isolated fp64 {a64 * 1e3:.1f} us, fp32 {a32 * 1e3:.1f} us; "
                      f"in-solve {'; '.join(line)}; graph loops {', '.join(loops)}", flush=True)
        fa.set_tuning(K, lattice=0, sell_map=0, time_kernels=0)
        for blocks in (256, 384, 512, 640, 1024, 2048):
            fa.set_tuning(K, lattice_blocks=blocks)
            a32 = fa.spmv_bench(K, mixed=True, reps=20)
            fa.set_tuning(K, lattice_blocks=blocks, time_kernels=1)
            r = fa.solve(K, f, mode="mixed", rtol=1e-10, download=False)
            t = fa.timings(K)
            fa.set_tuning(K, lattice_blocks=blocks, time_kernels=0)
            r = fa.solve(K, f, mode="mixed", rtol=1e-10, download=False)
            print(f"lattice blocks {blocks}: isolated fp32 {a32 * 1e3:.1f} us; in-solve "
                  f"{t['spmv_ms'] / max(t['spmv_launches'], 1) * 1e3:.1f} us; graph loop "
                  f"{fa.timings(K)['solve_ms']:.1f} ms ({r.iterations} its)", flush=True)
        fa.set_tuning(K, lattice_blocks=0)
    if args.cell_sweep:
        for march, rows, mname, counts in ((0, 2, "z-march, 2 rows/thread", (512, 640, 704, 768, 832, 896, 1024)),
                                           (0, 1, "z-march, 1 row/thread", (512, 768, 1024, 1536)),
                                           (-1, 0, "tiles", (512, 768, 1024, 1536))):
            for blocks in counts:
                smap = 0
                fa.set_tuning(K, cell_operator=1, cell_blocks=blocks, sell_map=smap, cell_march=march, cell_rows=rows)
                a64 = fa.spmv_bench(K, mixed=False, reps=20)
                a32 = fa.spmv_bench(K, mixed=True, reps=20)
                fa.set_tuning(K, cell_operator=1, cell_blocks=blocks, sell_map=smap, time_kernels=1)
                line = []
                for mode in ("cg", "mixed"):
                    r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
                    t = fa.timings(K)
                    line.append(f"{mode} [{r.operator}] {t['spmv_ms'] / max(t['spmv_launches'], 1) * 1e3:.1f} us x "
                                f"{t['spmv_launches']} ({r.iterations} its)")
                fa.set_tuning(K, cell_operator=1, cell_blocks=blocks, sell_map=smap, time_kernels=0)
                loops = []
                for mode in ("cg", "mixed"):
                    r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
                    loops.append(f"{mode} {fa.timings(K)['solve_ms']:.1f} ms")
                print(f"cell operator {mname} {blocks} blocks: isolated fp64 {a64 * 1e3:.1f} us, fp32 {a32 * 1e3:.1f} us; "
                      f"in-solve {'; '.join(line)}; graph loops {', '.join(loops)}", flush=True)
        fa.set_tuning(K, cell_operator=0, cell_blocks=0, sell_map=0, cell_march=0, cell_rows=0)
    if args.sell_sweep:
        for smap, mname, unrolls in ((1, "grid-stride, ordinary loads", (4, 8, 16)), (2, "contiguous/XCD", (4, 8)),
                                     (3, "grid-stride, non-temporal loads", (4, 8))):
            for unroll in unrolls:
                if smap == 3:
                    fa.set_tuning(K, sell_map=1, sell_nt=1, sell_unroll=unroll)
                else:
                    fa.set_tuning(K, sell_map=smap, sell_nt=-1, sell_unroll=unroll)
                a64 = fa.spmv_bench(K, mixed=False, reps=20)
                a32 = fa.spmv_bench(K, mixed=True, reps=20)
                fa.set_tuning(K, sell_map=smap if smap != 3 else 1, sell_unroll=unroll, time_kernels=1,
                              sell_nt=1 if smap == 3 else -1)
                line = []
                for mode in ("cg", "mixed"):
                    r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
                    t = fa.timings(K)
                    line.append(f"{mode} {t['spmv_ms'] / max(t['spmv_launches'], 1) * 1e3:.1f} us x "
                                f"{t['spmv_launches']} ({r.iterations} its, loop {t['solve_ms']:.1f} ms)")
                print(f"sell {mname} unroll {unroll}: isolated fp64 {a64 * 1e3:.1f} us, fp32 {a32 * 1e3:.1f} us; "
                      f"in-solve {'; '.join(line)}", flush=True)
        fa.set_tuning(K, sell_map=0, sell_unroll=0, time_kernels=0, sell_nt=0)
    if args.sweep:
        a64 = fa.spmv_bench(K, mixed=False, reps=20)
        a32 = fa.spmv_bench(K, mixed=True, reps=20)
        print(f"sweep spmv sell: fp64 {a64:.4f} ms, fp32+u16 {a32:.4f} ms", flush=True)
        fa.set_tuning(K, sell_compress=-1)
        a64 = fa.spmv_bench(K, mixed=False, reps=20)
        a32 = fa.spmv_bench(K, mixed=True, reps=20)
        print(f"sweep spmv sell int32 columns: fp64 {a64:.4f} ms, fp32 {a32:.4f} ms", flush=True)
        fa.set_tuning(K, sell_compress=0)
        for mode, name in ((2, "gather"), (1, "colored scatter")):
            fa.set_tuning(K, asm_mode=mode)
            ts = []
            for prec in ("fp64", "mixed"):
                fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
                fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
                ts.append(fa.timings(K)["assemble_kernel_ms"] + fa.timings(K)["fill_ms"])
            print(f"sweep assembly {name}: fp64 {ts[0]:.3f} ms, fp32-element {ts[1]:.3f} ms (incl. fill)",
                  flush=True)
        fa.set_tuning(K, asm_mode=1)
        for bs in (64, 128, 256):
            fa.set_tuning(K, asm_block=bs)
            ts = []
            for prec in ("fp64", "mixed"):
                fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
                ts.append(fa.timings(K)["assemble_kernel_ms"])
            print(f"sweep assembly block {bs}: fp64 {ts[0]:.3f} ms, fp32-element {ts[1]:.3f} ms", flush=True)
        for alg in ("greedy",):
            cols = fa.create_coloring(g, alg)
            fa.set_tuning(K, asm_block=0, asm_mode=1)
            fa.assemble_global(fa.start_assemble(K, f), cv, cols, el, precision="fp64")
            print(f"sweep coloring {alg} ({len(cols)} colors): fp64 {fa.timings(K)['assemble_kernel_ms']:.3f} ms",
                  flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
