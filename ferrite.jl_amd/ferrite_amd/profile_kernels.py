"""Kernel-level profiling driver for rocprofv3 (counters and per-kernel durations).

    python -m ferrite_amd.profile_kernels [--n 128] [--reps 5] [--solve]
    python -m ferrite_amd.profile_kernels --q2 [--n 48] [--reps 3]

Scalar Q1 (default): builds the 3D Poisson Q1 hex n³ system of the headline benchmark, then
runs, in order:
  1. a plain device copy of a known size (torch), the calibration for FETCH_SIZE/WRITE_SIZE;
  2. `reps` assemblies (FP64 element kernel), apply!;
  3. `reps` FP64 and FP32 solver SpMVs;
  4. with --solve: one FP64 and one mixed solve on their default operators, then the mixed
     solver on the lattice planes and the FP64 solver on the cell form.
--q2: builds the Q2 hex³ elasticity n³ system, runs the copy calibration, then `reps`
assemblies through the owner-computes path in FP32 (MFMA) and FP64 element precision, `reps`
solver SpMVs on the element matrices and on the sliced copy, and one colored-scatter assembly
of each precision.
Each phase prints a line, so a counter run can be matched to its kernels by name.
"""
import argparse
import os
import sys


def _copy_calibration(mib):
    import torch
    nbytes = mib << 20
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"phase copy: {nbytes} bytes read + {nbytes} bytes written", flush=True)
    # the measured copy rate (read + written bytes / time), the practical HBM peak
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(f"phase copy timing: {ms:.3f} ms per copy = {2 * nbytes / (ms * 1e-3) / 1e9:.0f} GB/s "
          f"(read + written)", flush=True)


def poisson(args):
    import numpy as np
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
    _copy_calibration(args.copy_mib)
    for _ in range(args.reps):
        fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="fp64")
    t = fa.timings(K)
    print(f"phase assemble: {args.reps} x {g.ncells} cells, last {t['assemble_kernel_ms']:.3f} ms "
          f"over {t['assemble_launches']} launches", flush=True)
    fa.apply(K, f, ch)
    ms64 = fa.spmv_bench(K, mixed=False, reps=args.reps)
    ms32 = fa.spmv_bench(K, mixed=True, reps=args.reps)
    print(f"phase spmv: fp64 {ms64:.4f} ms, fp32 {ms32:.4f} ms per launch (nnz {K.nnz}, n {dh.ndofs})",
          flush=True)
    if args.solve:
        for mode, tune in (("cg", {}), ("mixed", {}), ("mixed", {"cell_operator": -1}), ("cg", {"cell_operator": 1})):
            fa.set_tuning(K, **tune)
            r = fa.solve(K, f, mode=mode, rtol=1e-10, download=False)
            print(f"phase solve {mode} [{r.operator}]: {r.iterations} iterations, relres {r.relres:.2e}", flush=True)
        fa.set_tuning(K, cell_operator=0)


def q2(args):
    import numpy as np
    import ferrite_amd as fa
    n = args.n
    g = fa.generate_grid("Hexahedron", (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ip = fa.Lagrange("RefHexahedron", 2) ** 3
    cv = fa.CellValues(fa.QuadratureRule("RefHexahedron", 3), ip)
    dh = fa.DofHandler(g).add("u", ip).close()
    K = fa.allocate_matrix(fa.MI355XCSC, dh)
    f = fa.allocate_vector(K)
    colors = fa.create_coloring(g)
    el = fa.ElasticityElement(E=np.random.default_rng(0).uniform(0.5, 2.0, g.ncells), nu=0.3)
    _copy_calibration(args.copy_mib)
    for prec in ("mixed", "fp64"):
        for _ in range(args.reps):
            fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
        t = fa.timings(K)
        print(f"phase q2 gather {prec}: {args.reps} x {g.ncells} cells, last {t['assemble_kernel_ms']:.3f} ms "
              f"(element {t['element_ms']:.3f} ms, gather {t['scatter_ms']:.3f} ms)", flush=True)
    # solver SpMVs on the FP64-assembled K: the element matrices against the sliced copy
    fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision="fp64")
    for elem, name in ((1, "element matrices"), (-1, "sliced copy")):
        fa.set_tuning(K, elem_operator=elem)
        ms32 = fa.spmv_bench(K, mixed=True, reps=args.reps)
        ms64 = fa.spmv_bench(K, mixed=False, reps=args.reps)
        print(f"phase q2 spmv on the {name}: fp32 {ms32:.4f} ms, fp64 {ms64:.4f} ms per launch", flush=True)
    fa.set_tuning(K, elem_operator=0, asm_mode=1)
    for prec in ("mixed", "fp64"):
        fa.assemble_global(fa.start_assemble(K, f), cv, colors, el, precision=prec)
        t = fa.timings(K)
        print(f"phase q2 scatter {prec}: {t['assemble_kernel_ms']:.3f} ms over {t['assemble_launches']} launches",
              flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=0, help="cells per axis (default 128, 48 with --q2)")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--solve", action="store_true", help="also run the solves (scalar Q1)")
    ap.add_argument("--q2", action="store_true", help="the Q2 hex elasticity assembly instead")
    args = ap.parse_args()
    import torch  # noqa: F401  (torch's HIP runtime first: one runtime per process)
    if args.n == 0:
        args.n = 48 if args.q2 else 128
    (q2 if args.q2 else poisson)(args)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
