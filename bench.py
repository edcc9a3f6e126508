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


# what each method runs, in plain words (the bench line's config.method_label)
LABELS = {
    "mixed-learned-2l": "FP32 two-level PCG in FP64 refinement; learned blocks probed for 16 FP32 iterations "
                        "against block Jacobi, kept only if they win (learned_fallback = 1: rejected, block "
                        "Jacobi used)",
    "mixed-bj-2l": "FP32 two-level block-Jacobi PCG in FP64 refinement",
    "fp64-bj-2l": "FP64 two-level block-Jacobi PCG",
    "mixed-learned": "FP64 PCG to 1e-3, then FP32 learned-block PCG in FP64 refinement (probed)",
    "mixed-bj": "FP64 PCG to 1e-3, then FP32 block-Jacobi PCG in FP64 refinement",
    "mixed-bj-plain": "FP32 block-Jacobi PCG in FP64 refinement",
    "fp64-bj": "FP64 block-Jacobi PCG",
}
REFERENCE_FIXTURE = os.path.join(ROOT, "tests", "golden", "headline_160x16x16_seed1.npz")

# headline candidates: (jfm_solve method, preconditioner, extra options)
# (two-level: aggregates of about 1,536 nodes, 261 of them at 1.05M DOFs, and
# convergence checked every 16 iterations; both measured best for each
# method against 1,024-3,072 and 12-64)
METHODS = {
    "mixed-learned-2l": (1, 1, dict(learned_probe=16, inner_rtol=2e-3, coarse_nodes=1536, check_every=16)),
    "mixed-bj-2l": (1, 0, dict(inner_rtol=2e-3, coarse_nodes=1536, check_every=16)),
    "fp64-bj-2l": (0, 0, dict(coarse_nodes=1536, check_every=16)),
    "mixed-learned": (1, 1, dict(learned_probe=16, inner_rtol=1e-2, first_rtol=1e-3)),
    "mixed-bj": (1, 0, dict(inner_rtol=1e-2, first_rtol=1e-3)),
    "mixed-bj-plain": (1, 0, dict(inner_rtol=1e-4)),
    "fp64-bj": (0, 0, {}),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_matfree_fp32_tet10_160x16x16.json")


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
    if op in ("matfree", "ebe") and inf.get("n_patches", 0) > 0:
        # patch kernel: connectivity (int32 or 16-bit patch-relative, npe per
        # element; the EBE form reads int32), the geometric
        # factors (matrix-free) or the stored upper blocks of K_e (EBE), x
        # once, the patch-local incidences (uint16, npe per element), per
        # patch node its count (uint16) and target (int32); the fixed mask and
        # y of the interior nodes, the partial sums of the shared entries
        npn, nb, ns = inf["patch_nodes"], inf["shared_nodes"], inf["shared_slots"]
        interior = npn - ns
        rows = inf.get("patch_fp32_factor_rows", 11) if T == 4 else 11  # folded FP32 factors: 9
        el = rows * ngf * n_el * T if op == "matfree" else T * 9 * (npe * (npe + 1) // 2) * n_el
        cb = inf.get("patch_conn_bytes", 4) if op == "matfree" else 4
        patch = (cb * npe * n_el + el + 3 * N * T + 2 * npe * n_el + 6 * npn
                 + 3 * interior + 3 * interior * T + 3 * ns * T)
        # shared-node kernel: slot ranges and node ids, the partials, mask, y
        combine = 8 * nb + 4 + 3 * ns * T + 3 * nb + 3 * nb * T
        return {"patch": patch, "combine": combine}
    # element phase: connectivity + incidence slots (int32, npe each), the
    # geometric factors (11 per factor set) or the stored K_e blocks, x once,
    # the element results (3 npe per element) written once
    elem = 8 * npe * n_el + 11 * ngf * n_el * T + 3 * N * T + 3 * npe * n_el * T
    if op == "ebe":   # element matrices stored in the operator's precision
        elem = 8 * npe * n_el + T * 9 * (npe * (npe + 1) // 2) * n_el + 3 * N * T + 3 * npe * n_el * T
    # gather: node pointers, the element results read once, y and the fixed mask
    gather = 4 * (N + 1) + 3 * npe * n_el * T + 3 * N * T + 3 * N
    return {"element": elem, "gather": gather}


def ideal_ku(inf, prec):
    """Algorithmic work of one matrix-free K.u by the per-unit figures of the
    hot-path measurement plan: bytes = 16n + 40 n_el + 24N + 8 n_el (FP64;
    coordinates, connectivity, E, x in, y out; no scratch) and 1.6 kflop per
    affine Tet10 element."""
    n, N, n_el = inf["n_dofs"], inf["n_nodes"], inf["n_el"]
    scale = 1.0 if prec == 0 else 0.5
    return {"bytes": int(scale * (16 * n + 24 * N + 8 * n_el) + 40 * n_el), "flops": 1600 * n_el}


def ebe_algorithmic(inf, prec):
    """Per-unit figure of the EBE K.u: symmetric-packed K_e (465 values per
    Tet10), int32 connectivity, x in and y out: FP64 3,720 n_el + 16n + 40 n_el,
    FP32 1,860 n_el + 8n + 40 n_el (bytes)."""
    T = 8 if prec == 0 else 4
    npe, n_el, n = inf["npe"], inf["n_el"], inf["n_dofs"]
    packed = (3 * npe) * (3 * npe + 1) // 2
    return T * packed * n_el + 2 * T * n + 4 * npe * n_el


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
    ap.add_argument("--skip-renumbering", action="store_true",
                    help="skip the K.u timing on shuffled and RCM-renumbered copies of the mesh")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the K.u kernels and known-size copies only (for counter runs)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="run only the node features and the network inference (for MFMA counter runs: "
                         "no PCG, no graph launches)")
    ap.add_argument("--workload", default="linear", choices=("linear", "nonlinear"),
                    help="linear: the 1.05M-DOF solve (default); nonlinear: the 8.1M-DOF neo-Hookean Newton")
    ap.add_argument("--method", default="mixed-learned-2l", choices=sorted(METHODS),
                    help="headline solve method")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "nonlinear":
        return nonlinear(args)

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
    if args.mlp_only:
        from jfem.learned import DEFAULT_WEIGHTS
        with open(DEFAULT_WEIGHTS, "rb") as fh:
            ctx.set_learned_weights(fh.read())
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            ctx.learned_outputs()       # features, network inference and blocks, copied back
            ts.append(time.perf_counter() - t0)
        print(json.dumps({"mlp_only": True, "n_nodes": inf["n_nodes"], "host_s": ts}), flush=True)
        return
    t0 = time.perf_counter()
    ctx.assemble(want_vals=False, want_f=False)
    t_asm = time.perf_counter() - t0
    log(f"assembly (K_e + CSR values + f) {t_asm * 1e3:.1f} ms host wall")

    # ---- K.u kernels --------------------------------------------------------
    kernels = {}
    for op in ("csr", "matfree", "ebe"):
        for prec in (0, 1):
            tot, k1, k2 = ctx.time_apply(op=op, precision=prec, reps=50)
            inf = ctx.info()  # the operator's structure exists once it has run
            b = kernel_bytes(inf, op, prec)
            name = f"{op}_{'fp64' if prec == 0 else 'fp32'}"
            ent = {"ms": tot}
            if op == "csr":
                ent["spmv_ms"] = k1
                ent["spmv_GBps"] = b["spmv"] / (k1 * 1e-3) / 1e9
            elif "patch" in b:
                ent["patch_ms"], ent["combine_ms"] = k1, k2
                ent["patch_GBps"] = b["patch"] / (k1 * 1e-3) / 1e9
                ent["combine_GBps"] = b["combine"] / (k2 * 1e-3) / 1e9
                ent["Ku_GBps"] = (b["patch"] + b["combine"]) / (tot * 1e-3) / 1e9
                if op == "matfree":
                    ent["patch_TFLOPs"] = flops_matfree_elem(inf) / (k1 * 1e-3) / 1e12
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

    renum = None if (args.skip_renumbering or args.kernels_only) else renumbering_ku(jfem, mesh, kernels)

    if args.kernels_only:
        copy = copy_calibration(ctx, inf["n_dofs"])
        print(json.dumps({"kernels": kernels, "copy": copy, "n_el": inf["n_el"], "n_dofs": inf["n_dofs"]}), flush=True)
        return

    # ---- solves ----------------------------------------------------------------
    method, precond, extra = METHODS[args.method]
    from jfem.learned import DEFAULT_WEIGHTS
    if precond == 1:
        with open(DEFAULT_WEIGHTS, "rb") as fh:
            ctx.set_learned_weights(fh.read())

    def solve(method, op, op_outer="matfree", precond=0, **kw):
        t = time.perf_counter()
        u, _, st = ctx.solve(method=method, op=op, op_outer=op_outer, precond=precond, rtol=1e-8,
                             raise_on_noconv=False, **{"inner_rtol": 1e-4, "check_every": 32, **kw})
        st["wall_s"] = time.perf_counter() - t
        return u, st

    def repeat(n, *a, **kw):
        """warm once, then n runs; (last u, stats of the median run, all times)"""
        solve(*a, **kw)
        runs = [solve(*a, **kw) for _ in range(n)]
        runs.sort(key=lambda r: r[1]["solve_ms"])
        u_, st_ = runs[len(runs) // 2]
        return u_, st_, [r[1]["solve_ms"] / 1e3 for r in runs]

    u64, st64, t64 = repeat(3, 0, "matfree")
    log(f"FP64 BJ-PCG matfree (median of 3 after a warm run): {st64}")
    _, st64c, t64c = repeat(3, 0, "csr", "csr")
    log(f"FP64 BJ-PCG csr (median of 3 after a warm run): {st64c}")
    others = {}
    for name in ("fp64-bj-2l", "mixed-bj-2l"):
        if name == args.method:
            continue
        m_, p_, kw_ = METHODS[name]
        _, st_, ts_ = repeat(3, m_, "matfree", precond=p_, **kw_)
        others[name] = {"s": st_["solve_ms"] / 1e3, "setup_s": st_["setup_ms"] / 1e3, "runs_s": ts_,
                        "iters": st_["iterations"], "relres": st_["relres"]}
        log(f"{name}: {st_}")

    blob = None
    if precond == 1:
        with open(DEFAULT_WEIGHTS, "rb") as fh:
            blob = fh.read()

    def headline():
        # setting the weights drops the cached learned blocks, so every timed
        # solve recomputes the node features and the network's blocks
        if blob is not None:
            ctx.set_learned_weights(blob)
        return solve(method, "matfree", precond=precond, **extra)

    for _ in range(args.warmup):
        headline()
    times, walls, stats = [], [], None
    u_head = None
    for _ in range(args.steps):
        u_head, st = headline()
        # solve_ms spans the whole call: preconditioner setup (coarse level,
        # node features, network inference and the probe) included
        times.append(st["solve_ms"] / 1e3)
        walls.append(st["wall_s"])
        stats = st
        log(f"{args.method} step: {st}")
    value = statistics.median(times)
    err_vs_fp64 = float(np.linalg.norm(u_head - u64) / np.linalg.norm(u64))
    err_oracle = error_vs_reference(mesh, u_head) if (nx, ny, nz, args.seed) == (160, 16, 16, 1) else None
    log(f"error vs the oracle's FP64 answer: {err_oracle}")

    # ---- roofline of the dominant kernel of the headline solve ---------------
    hk = kernels["matfree_fp32" if method == 1 else "matfree_fp64"]
    tname = "float" if method == 1 else "double"
    prec = 1 if method == 1 else 0
    sv = ideal_ku(inf, prec)
    if "patch_ms" in hk:
        c16 = ",C16" if inf.get("patch_conn_bytes", 4) == 2 else ""  # 16-bit patch-relative connectivity
        if method == 1 and inf.get("patch_fp32_factor_rows", 11) == 9:
            c16 += ",F9"  # folded FP32 factors, 9 rows
        dom, dom_ms, impl_bytes = f"k_ku_patch<10,{tname},AFF{c16}>", hk["patch_ms"], hk["bytes"]["patch"]
    elif hk["element_ms"] >= hk["gather_ms"]:
        dom, dom_ms, impl_bytes = f"k_mf_elem<10,{tname},AFF>", hk["element_ms"], hk["bytes"]["element"]
    else:
        dom, dom_ms, impl_bytes = f"k_gather<{tname}>", hk["gather_ms"], hk["bytes"]["gather"]
    # algorithmic bytes of the matrix-free K.u this kernel performs (every
    # element product and the nodal sums of the patch-interior nodes), per the
    # per-unit figure: coordinates, E, int32 connectivity, x in, y out
    algo = sv["bytes"]
    achieved = algo / (dom_ms * 1e-3) / 1e9
    peak_tf = FP32_PEAK_TFLOPS if method == 1 else FP64_PEAK_TFLOPS
    traffic = measured_traffic(dom, inf)
    ku_ms = hk["ms"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "bytes_per_launch": algo, "avg_launch_ms": dom_ms,
                "bytes_rule": "8n + 44 n_el + 12N (FP32: x in and y out 4 B per DOF, int32 connectivity 40 B "
                              "and E 4 B per element, coordinates 12 B per node)" if method == 1 else
                              "16n + 48 n_el + 24N (FP64)",
                "traffic_over_algorithmic": None if traffic is None else traffic / algo,
                "flops_per_launch": sv["flops"],
                "flop_frac": sv["flops"] / (dom_ms * 1e-3) / 1e12 / peak_tf,
                "implementation_bytes": impl_bytes,
                "implementation_GBps": impl_bytes / (dom_ms * 1e-3) / 1e9,
                "ku_whole": {"ms": ku_ms, "GBps": algo / (ku_ms * 1e-3) / 1e9,
                             "frac": algo / (ku_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                "ebe": {}}
    for key, pr in (("fp64", 0), ("fp32", 1)):
        ke = kernels.get(f"ebe_{key}")
        if ke:
            b = ebe_algorithmic(inf, pr)
            tn = 'double' if pr == 0 else 'float'
            # the patch form (stored-matrix element products summed in LDS,
            # then the shared-node kernel), or the two-phase one without patches
            k1 = ke.get("patch_ms", ke.get("element_ms"))
            roofline["ebe"][key] = {"kernel": f"k_ku_patch<10,{tn},EBE>" if "patch_ms" in ke else f"k_ebe_elem<10,{tn}>",
                                   "bytes_per_launch": b, "element_ms": k1, "ku_ms": ke["ms"],
                                   "element_frac": b / (k1 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "ku_frac": b / (ke["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}

    # ---- CPU baseline (oracle, bounded samples) -------------------------------
    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(mesh, st64["iterations"], args.cpu_iters)

    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-inner/f64-outer" if method == 1 else "f64",
        "data": "synthetic (seeded jittered box mesh)",
        "config": {"workload": f"Tet10 cantilever {nx}x{ny}x{nz} hexes x6, {inf['n_dofs']} DOFs, "
                               f"log-normal E sigma 0.5, rtol 1e-8", "n_dofs": inf["n_dofs"],
                   "nnz": inf["nnz"], "method": args.method, "method_label": LABELS[args.method],
                   "method_options": extra, "timed": "whole jfm_solve call from device-resident mesh and load: "
                   "coarse-level setup, node features, network inference, probe, PCG and FP64 refinement",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_pcg_s": st64["solve_ms"] / 1e3, "fp64_pcg_runs_s": t64, "fp64_pcg_iters": st64["iterations"],
        "fp64_pcg_csr_s": st64c["solve_ms"] / 1e3, "fp64_pcg_csr_runs_s": t64c,
        "fp64_pcg_csr_iters": st64c["iterations"],
        "speedup_vs_fp64_matfree": st64["solve_ms"] / 1e3 / value,
        "speedup_vs_fp64_csr": st64c["solve_ms"] / 1e3 / value,
        "speedup_vs_fp64_2l": (others["fp64-bj-2l"]["s"] / value) if "fp64-bj-2l" in others else None,
        "iters": stats["iterations"], "outer": stats["outer"], "relres": stats["relres"],
        "learned_fallback": stats["fallback"], "setup_ms": stats["setup_ms"], "other_methods": others,
        "err_vs_oracle_l2": err_oracle, "err_vs_fp64_pcg_l2": err_vs_fp64,
        "host_wall_s": statistics.median(walls),
        "dofs_per_s": inf["n_dofs"] / value, "upload_s": t_upload, "assembly_wall_s": t_asm,
        "kernels": kernels, "ku_renumbering": renum,
    }
    print(json.dumps(out), flush=True)


def renumbering_ku(jfem, mesh, kernels):
    """Matrix-free K.u time of the same mesh with its node ids shuffled, and
    after reverse Cuthill-McKee renumbering of the shuffled ids (the library's
    own element order, which follows the node ids, changes with them)."""
    import numpy as np
    from jfem.mesh import node_bandwidth, rcm_order, renumber
    out = {"natural": {"bandwidth": node_bandwidth(mesh),
                       "fp32_ms": kernels["matfree_fp32"]["ms"], "fp64_ms": kernels["matfree_fp64"]["ms"]}}
    sh = renumber(mesh, np.random.default_rng(12).permutation(mesh.n_nodes))
    t0 = time.perf_counter()
    perm = rcm_order(sh)
    t_rcm = time.perf_counter() - t0
    for name, mm in (("shuffled", sh), ("rcm", renumber(sh, perm))):
        c = jfem.Context(0)
        c.upload(mm)
        ent = {"bandwidth": node_bandwidth(mm)}
        for prec, key in ((1, "fp32_ms"), (0, "fp64_ms")):
            ent[key] = c.time_apply(op="matfree", precision=prec, reps=50)[0]
        c.close()
        out[name] = ent
    out["rcm_host_s"] = t_rcm
    log(f"K.u by numbering: {out}")
    return out


def error_vs_reference(mesh, u):
    """Relative L2 error of u against the oracle's FP64 answer for this mesh
    (tests/golden/headline_160x16x16_seed1.npz): the larger of the error over
    every 16th DOF and the estimate from 16 random +-1 projections."""
    import numpy as np
    from jfem.mesh import splitmix_uniform
    try:
        z = np.load(REFERENCE_FIXTURE)
    except OSError:
        return None
    s = int(z["idx_stride"])
    e_sample = np.linalg.norm(u[::s] - z["u_sample"]) / np.linalg.norm(z["u_sample"])
    proj = np.array([np.where(splitmix_uniform(1, 40 + k, np.arange(u.size)) < 0.5, -1.0, 1.0) @ u
                     for k in range(z["proj"].size)])
    e_proj = np.sqrt(np.mean((proj - z["proj"]) ** 2)) / float(z["norm2"])
    return float(max(e_sample, e_proj))


NL_TRACTION = 52.5   # end traction giving a tip deflection of about 10% of the span


def nonlinear(args):
    """8.1M-DOF neo-Hookean cantilever: 5 load steps of Newton to FP64 1e-8,
    tangent solves by FP32 matrix-free PCG (learned blocks with the runtime
    probe, two-level) inside FP64 refinement.  One step = the whole
    load-stepped Newton solve."""
    import numpy as np
    import jfem
    from jfem.mesh import box_tet10
    nx, ny, nz = (320, 32, 32) if args.grid == [160, 16, 16] else args.grid
    t0 = time.perf_counter()
    mesh = box_tet10(nx, ny, nz, traction=(0.0, 0.0, -NL_TRACTION))
    log(f"mesh {nx}x{ny}x{nz}: {mesh.n_nodes} nodes, {mesh.n_el} Tet10, {mesh.n_dofs} DOFs "
        f"({time.perf_counter() - t0:.2f}s)")
    ctx = jfem.Context(0)
    t0 = time.perf_counter()
    ctx.upload(mesh)
    inf = ctx.info()
    log(f"upload {time.perf_counter() - t0:.2f}s, device {inf['device_bytes'] / 2**30:.2f} GiB")
    from jfem.learned import DEFAULT_WEIGHTS
    with open(DEFAULT_WEIGHTS, "rb") as fh:
        ctx.set_learned_weights(fh.read())
    opts = dict(load_steps=5, rtol=1e-8, method=1, precond=1, learned_probe=16, inner_rtol=1e-2,
                coarse_nodes=4096, max_newton=20)
    times, stats, u = [], None, None
    for k in range(args.warmup + args.steps):
        t = time.perf_counter()
        u, st = ctx.solve_nonlinear(**opts)
        st["wall_s"] = time.perf_counter() - t
        log(f"{'warmup' if k < args.warmup else 'step'}: {st}")
        if k >= args.warmup:
            times.append(st["solve_ms"] / 1e3)
            stats = st
    value = statistics.median(times)
    U = u.reshape(-1, 3)
    tip = U[mesh.X[:, 0] > 10.0 - 1e-9].mean(axis=0)
    t_apply = ctx.time_apply(op="tangent", precision=1, reps=20)
    # the linear matrix-free K.u at this size, where the launch latency that
    # bounds it at 1M DOFs is spread over 8x the work
    ku_large = {}
    for prec in (0, 1):
        tot, k1, k2 = ctx.time_apply(op="matfree", precision=prec, reps=20)
        inf2 = ctx.info()
        b = kernel_bytes(inf2, "matfree", prec)
        (n1, b1), (n2, b2) = list(b.items())   # patch/combine, or element/gather
        sv = ideal_ku(inf2, prec)
        ku_large["fp64" if prec == 0 else "fp32"] = {
            "ms": tot, f"{n1}_ms": k1, f"{n2}_ms": k2,
            f"{n1}_GBps": b1 / (k1 * 1e-3) / 1e9, f"{n1}_frac": b1 / (k1 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "Ku_GBps": (b1 + b2) / (tot * 1e-3) / 1e9,
            "ku_ideal_GBps": sv["bytes"] / (tot * 1e-3) / 1e9, "bytes": b}
        log(f"K.u {'fp64' if prec == 0 else 'fp32'} at {inf2['n_dofs']} DOFs: {ku_large['fp64' if prec == 0 else 'fp32']}")
    cpu = None
    if not args.skip_cpu:
        cpu = nl_cpu_baseline(mesh, u, stats["newton_iterations"])
    out = {
        "metric": "8.1M-DOF neo-Hookean Newton solve time @res 1e-8 (5 load steps)", "value": value, "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": value * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32-inner/f64-outer",
        "data": "synthetic (seeded box mesh)",
        "config": {"workload": f"neo-Hookean Tet10 cantilever {nx}x{ny}x{nz} hexes x6, {inf['n_dofs']} DOFs, "
                               f"mu 80.77e3, lambda 121.15e3, end traction {NL_TRACTION}", "n_dofs": inf["n_dofs"],
                   "options": opts, "parallelism": "single GPU"},
        "newton_iterations": stats["newton_iterations"], "linear_iterations": stats["linear_iterations"],
        "learned_fallbacks": stats["fallbacks"], "relres": stats["relres"], "converged": stats["converged"],
        "tip_displacement": tip.tolist(), "tangent_apply_fp32_ms": t_apply[0], "linear_ku": ku_large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def nl_cpu_baseline(mesh, u, newton_iterations, sample_el=20000):
    """Oracle TL residual + tangent for a bounded sample of elements, scaled to
    the whole mesh and to the Newton iterations the GPU took.  A lower bound of
    the CPU path: it leaves out the tangent solves themselves."""
    import numpy as np
    from oracle import oracle as O
    from jfem.mesh import Mesh
    k = min(sample_el, mesh.n_el)
    sub = Mesh(X=mesh.X, conn=mesh.conn[:k], etype=mesh.etype, E=mesh.E[:k], nu=mesh.nu[:k],
               faces=mesh.faces[:0], ftype=mesh.ftype, clamp_nodes=mesh.clamp_nodes)
    t0 = time.perf_counter()
    O.nh_assemble(sub, u, tangent=True)
    dt = time.perf_counter() - t0
    value = dt * mesh.n_el / k * newton_iterations
    log(f"cpu oracle: TL residual+tangent of {k} elements {dt:.2f}s -> {value:.1f}s for {newton_iterations} Newton iterations")
    return {"value": value, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle TL residual and tangent (COO) of {k} of {mesh.n_el} elements ({dt:.2f} s, one core), "
                      f"scaled to the mesh and to {newton_iterations} Newton iterations; tangent solves not included"}


def copy_calibration(ctx, n_dofs):
    """Device-to-device copies of known size for scaling the FETCH_SIZE and
    WRITE_SIZE counters (rocprofv3 --pmc) to bytes: a few PCG starts, each of
    which copies FP64 vectors of n_dofs entries with hipMemcpyAsync (the
    runtime's copyBuffer kernel)."""
    for _ in range(3):
        ctx.solve(method=0, op="matfree", rtol=1e-30, maxit=2, raise_on_noconv=False)
    return {"kernel": "__amd_rocclr_copyBuffer", "bytes_read": 8 * n_dofs, "bytes_written": 8 * n_dofs}


def measured_traffic(kernel, inf):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled), when it covers this mesh."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or t.get("n_el") != inf["n_el"]:
        return None
    return k["bytes_per_launch"]


def host_cpu():
    """(model name, usable cores) of the host running the oracle."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, cores


def _median_time(fn, reps):
    ts = []
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), ts, out


def cpu_baseline(mesh, iters_needed, sample_iters):
    """The oracle (the reference's algorithm restated in C/OpenMP, and LAPACK
    for the direct solves) on the host, each figure the median of 3 runs:
      * config-1 COO->CSR assembly, all cores;
      * config-1 FP64 block-Jacobi PCG iterations (a sample of `sample_iters`
        on all cores, 5 on one core), extrapolated to the `iters_needed`
        iterations the method takes;
      * the reference's own solver path, a Dirichlet-eliminated direct solve
        (banded Cholesky, LAPACK), on the 40x4x4 mesh (all cores and one core)
        and on the 80x8x8 jittered mesh (all cores; one run).
    `value` is the config-1 time-to-solution on all cores: assembly + PCG."""
    import numpy as np
    from oracle import oracle as O
    from jfem.mesh import box_tet10
    model, avail = host_cpu()
    cores = min(avail, 16)
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        threadpool_limits = None
    t_asm, asm_runs, csr = _median_time(lambda: O.assemble_csr(mesh, nthreads=cores), 3)
    f = O.load_vector(mesh)
    fixed = mesh.fixed_mask()

    def pcg(k, nt):
        return O.pcg_bj(csr, fixed, f, rtol=1e-30, maxit=k, nthreads=nt)

    t_pcg, _, _ = _median_time(lambda: pcg(sample_iters, cores), 3)
    t_it = t_pcg / sample_iters
    t_pcg1, _, _ = _median_time(lambda: pcg(5, 1), 3)
    t_it1 = t_pcg1 / 5
    del csr
    rows = {"cpu_model": model, "cores_available": avail,
            "config1_assembly_s": {"cores": cores, "median_s": t_asm, "runs_s": asm_runs},
            "config1_pcg_iteration_s": {str(cores): t_it, "1": t_it1}}
    log(f"cpu oracle ({model}, {cores} of {avail} cores): assembly {t_asm:.2f}s, PCG {t_it * 1e3:.1f} ms/iter "
        f"({cores} cores), {t_it1 * 1e3:.1f} ms/iter (1 core)")
    # the direct solves
    for name, m in (("40x4x4", box_tet10(40, 4, 4)), ("80x8x8", box_tet10(80, 8, 8, seed=7, jitter=0.1, E_sigma=0.5))):
        c = O.assemble_csr(m, nthreads=cores)
        fm = O.load_vector(m)
        g = np.zeros(m.n_dofs)
        reps = 3 if name == "40x4x4" else 1
        td, runs, _ = _median_time(lambda: O.solve_direct_lapack(c, m.fixed_mask(), g, fm), reps)
        ent = {"n_dofs": m.n_dofs, "median_s": td, "runs_s": runs, "cores": cores}
        if name == "40x4x4" and threadpool_limits is not None:
            with threadpool_limits(limits=1):
                td1, runs1, _ = _median_time(lambda: O.solve_direct_lapack(c, m.fixed_mask(), g, fm), 3)
            ent["one_core_median_s"] = td1
            ent["one_core_runs_s"] = runs1
        rows[f"direct_solve_{name}"] = ent
        log(f"cpu oracle direct solve {name}: {td:.2f}s ({ent})")
        del c
    value = t_asm + t_it * iters_needed
    return {"value": value, "unit": "s", "cores": cores, "kind": "port",
            "sample": f"oracle COO->CSR assembly of the full mesh (median of 3: {t_asm:.1f} s) + FP64 block-Jacobi "
                      f"PCG ({sample_iters} iterations, median of 3: {t_it * 1e3:.1f} ms each, OpenMP) "
                      f"extrapolated to the {iters_needed} iterations the method needs; one core: "
                      f"{t_it1 * 1e3:.1f} ms per iteration; host {model}",
            "rows": rows}


if __name__ == "__main__":
    main()
