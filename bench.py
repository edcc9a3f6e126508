"""Benchmark: 1.05M-DOF Tet10 cantilever solved to FP64 relative residual 1e-8.

One step = one solve of K u = f (free DOFs) from resident device data with the
headline method (FP32 matrix-free PCG inside FP64 iterative refinement),
timed on the device (HIP events inside the library) and on the host clock.
Printed beside it: the plain FP64 block-Jacobi PCG on the same GPU, the K.u
kernels' HBM roofline, and the oracle's CPU time on a bounded sample.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "juliafem.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "1M-DOF Tet10 solve time @res 1e-8: learned+FP32 vs FP64 PCG; K·u GB/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(inf, op, prec):
    """Algorithmic HBM bytes per launch of each K.u kernel (unique data touched once)."""
    T = 8 if prec == 0 else 4
    N, n_el, npe, nnzb = inf["n_nodes"], inf["n_el"], inf["npe"], inf["nnzb"]
    ngf = 1 if inf["affine"] else {10: 4, 4: 1, 8: 8}[npe]
    if op == "csr":
        spmv = 9 * nnzb * T + 4 * nnzb + 4 * (N + 1) + 2 * 3 * N * T + 3 * N
        return {"spmv": spmv}
    elem = 4 * npe * n_el + 11 * ngf * n_el * T + 3 * N * T + 3 * npe * n_el * T
    if op == "ebe":
        elem = 4 * npe * n_el + 8 * 9 * (npe * (npe + 1) // 2) * n_el + 3 * N * T + 3 * npe * n_el * T
    gather = 4 * (N + 1) + 4 * npe * n_el + 3 * npe * n_el * T + 3 * N * T + 3 * N
    return {"element": elem, "gather": gather}


def flops_matfree_elem(inf):
    # per Gauss point: reference gradient 2*27 (non-zero dN only ~ 2*3*27),
    # grad = H invJ' 2*27, stress ~12, P = invJ' sigma 2*27, nodal forces 2*3*27
    per_gp = 2 * 81 + 2 * 27 + 12 + 2 * 27 + 2 * 81
    return inf["n_el"] * 4 * per_gp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, nargs=3, default=[160, 16, 16])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import jfem
    from jfem.mesh import box_tet10

    nx, ny, nz = args.grid
    t0 = time.perf_counter()
    mesh = box_tet10(nx, ny, nz, seed=args.seed, jitter=0.1, E_sigma=0.5)
    log(f"mesh {nx}x{ny}x{nz}: {mesh.n_nodes} nodes, {mesh.n_el} Tet10, {mesh.n_dofs} DOFs "
        f"({time.perf_counter() - t0:.2f}s)")
    ctx = jfem.Context(0)
    t0 = time.perf_counter()
    ctx.upload(mesh)
    t_upload = time.perf_counter() - t0
    inf = ctx.info()
    log(f"upload+pattern {t_upload:.2f}s nnz={inf['nnz']} affine={inf['affine']}")
    t0 = time.perf_counter()
    ctx.assemble(want_vals=False, want_f=False)
    t_asm = time.perf_counter() - t0
    log(f"assembly (K_e + CSR values + f) {t_asm * 1e3:.1f} ms host wall")

    # ---- K.u kernels --------------------------------------------------------
    kernels = {}
    for op in ("csr", "matfree", "ebe"):
        for prec in (0, 1):
            if op == "ebe" and prec == 1:
                continue
            tot, k1, k2 = ctx.time_apply(op=op, precision=prec, reps=50)
            b = kernel_bytes(inf, op, prec)
            name = f"{op}_{'fp64' if prec == 0 else 'fp32'}"
            ent = {"ms": tot}
            if op == "csr":
                ent["spmv_ms"] = k1
                ent["spmv_GBps"] = b["spmv"] / (k1 * 1e-3) / 1e9
            else:
                ent["element_ms"], ent["gather_ms"] = k1, k2
                ent["element_GBps"] = b["element"] / (k1 * 1e-3) / 1e9
                ent["gather_GBps"] = b["gather"] / (k2 * 1e-3) / 1e9
                ent["Ku_GBps"] = (b["element"] + b["gather"]) / (tot * 1e-3) / 1e9
                if op == "matfree":
                    ent["element_TFLOPs"] = flops_matfree_elem(inf) / (k1 * 1e-3) / 1e12
            ent["bytes"] = b
            kernels[name] = ent
            log(f"K.u {name}: {json.dumps({k: (round(v, 4) if isinstance(v, float) else v) for k, v in ent.items()})}")

    # ---- solves ----------------------------------------------------------------
    def solve(method, op, op_outer="matfree"):
        t = time.perf_counter()
        u, _, st = ctx.solve(method=method, op=op, op_outer=op_outer, rtol=1e-8, inner_rtol=1e-4,
                             check_every=32, raise_on_noconv=False)
        st["wall_s"] = time.perf_counter() - t
        return u, st

    u64, st64 = solve(0, "matfree")
    log(f"FP64 BJ-PCG matfree: {st64}")
    u64c, st64c = solve(0, "csr", "csr")
    log(f"FP64 BJ-PCG csr: {st64c}")

    for _ in range(args.warmup):
        solve(1, "matfree")
    times, walls, stats = [], [], None
    u_mixed = None
    for _ in range(args.steps):
        u_mixed, st = solve(1, "matfree")
        times.append(st["solve_ms"] / 1e3)
        walls.append(st["wall_s"])
        stats = st
        log(f"mixed step: {st}")
    value = statistics.median(times)
    err_vs_fp64 = float(np.linalg.norm(u_mixed - u64) / np.linalg.norm(u64))

    # ---- roofline of the dominant kernel of the headline solve ---------------
    mf32 = kernels["matfree_fp32"]
    if mf32["element_ms"] >= mf32["gather_ms"]:
        dom, dom_ms, dom_bytes = "k_mf_elem<10,float,true>", mf32["element_ms"], mf32["bytes"]["element"]
    else:
        dom, dom_ms, dom_bytes = "k_gather<float>", mf32["gather_ms"], mf32["bytes"]["gather"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms}

    # ---- CPU baseline (oracle, bounded sample) --------------------------------
    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(mesh, st64["iterations"], args.cpu_iters)

    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-inner/f64-outer", "data": "synthetic (seeded jittered box mesh)",
        "config": {"workload": f"Tet10 cantilever {nx}x{ny}x{nz} hexes x6, {inf['n_dofs']} DOFs, "
                               f"log-normal E sigma 0.5, rtol 1e-8", "n_dofs": inf["n_dofs"],
                   "nnz": inf["nnz"], "method": "FP32 matrix-free block-Jacobi PCG (inner rtol 1e-4) in FP64 "
                                                 "iterative refinement", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_pcg_s": st64["solve_ms"] / 1e3, "fp64_pcg_iters": st64["iterations"],
        "fp64_pcg_csr_s": st64c["solve_ms"] / 1e3, "fp64_pcg_csr_iters": st64c["iterations"],
        "speedup_vs_fp64_matfree": st64["solve_ms"] / 1e3 / value,
        "speedup_vs_fp64_csr": st64c["solve_ms"] / 1e3 / value,
        "mixed_iters": stats["iterations"], "mixed_outer": stats["outer"], "relres": stats["relres"],
        "err_vs_fp64_l2": err_vs_fp64, "host_wall_s": statistics.median(walls),
        "dofs_per_s": inf["n_dofs"] / value, "upload_s": t_upload, "assembly_wall_s": t_asm,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(mesh, iters_needed, sample_iters):
    """Oracle on the host: full COO->CSR assembly of the same mesh, then
    `sample_iters` FP64 block-Jacobi PCG iterations; the time-to-solution is
    assembly + per-iteration time x the iteration count the same method needs."""
    import numpy as np
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, 16)
    t0 = time.perf_counter()
    csr = O.assemble_csr(mesh, nthreads=cores)
    t_asm = time.perf_counter() - t0
    f = O.load_vector(mesh)
    t0 = time.perf_counter()
    _, it, rel, _ = O.pcg_bj(csr, mesh.fixed_mask(), f, rtol=1e-30, maxit=sample_iters, nthreads=cores)
    t_it = (time.perf_counter() - t0) / max(it, 1)
    del csr
    value = t_asm + t_it * iters_needed
    log(f"cpu oracle: assembly {t_asm:.1f}s, {t_it * 1e3:.1f} ms/iter x {iters_needed} iters")
    return {"value": value, "unit": "s", "cores": cores, "kind": "port",
            "sample": f"oracle COO->CSR assembly of the full mesh ({t_asm:.1f} s) + {it} FP64 block-Jacobi "
                      f"PCG iterations ({t_it * 1e3:.1f} ms each, OpenMP), extrapolated to the "
                      f"{iters_needed} iterations the method needs"}


if __name__ == "__main__":
    main()
