"""Train the learned block preconditioner and evaluate it.

    python -m jfem.train_learned [--meshes 64] [--steps 300] [--eval 8] [--coarse-nodes 512] [--out PATH]

(run from juliafem.jl_amd/, on a GPU).  Training data: seeded Tet10
cantilever meshes from 40x4x4 to 80x8x8 hexes, log-normal E with sigma drawn
from [0.25, 1.0], +-0.1 h interior jitter; residual/error pairs from the
iterates of the FP64 block-Jacobi PCG, two-level when --coarse-nodes > 0, in
which case the blocks are trained as the fine level of that preconditioner.
Evaluation: held-out seeds at 160x16x16 (1,048,707 DOFs), to 1e-8, FP64 PCG
and FP32 inner PCG in FP64 refinement, each with block Jacobi and with the
learned blocks (and the coarse level when --coarse-nodes > 0).
Writes the BF16 blob to --out and a JSON record (losses, iteration counts,
times) beside it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib
from .learned import DEFAULT_WEIGHTS, LearnedPreconditioner, harvest_pairs, train
from .mesh import box_tet10


def log(msg):
    print(f"[train] {msg}", flush=True)


def coarse_nodes_for(grid, ratio: float, L=(10.0, 1.0, 1.0)) -> int:
    """Aggregate size giving cells about `ratio` hex cells wide on this grid,
    so that training meshes see the coarse level the evaluation mesh sees."""
    nx, ny, nz = grid
    n_nodes = (2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1)
    h = L[0] / nx
    return max(8, int(round(n_nodes * (ratio * h) ** 3 / (L[0] * L[1] * L[2]))))


def training_meshes(n: int, seed0: int = 1000):
    rng = np.random.default_rng(seed0)
    out = []
    for k in range(n):
        scale = (1.0, 1.5, 2.0)[k % 3]     # 40x4x4, 60x6x6, 80x8x8
        nx, ny = int(40 * scale), int(4 * scale)
        sigma = float(rng.uniform(0.25, 1.0))
        out.append(dict(grid=(nx, ny, ny), seed=seed0 + k, sigma=sigma))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=64)
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--eval-grid", type=int, nargs=3, default=[160, 16, 16])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--coarse-nodes", type=int, default=2048,
                    help="coarse_nodes of the evaluation mesh (0: single-level training)")
    ap.add_argument("--coarse-ratio", type=float, default=6.25,
                    help="aggregate width in hex cells for the training meshes")
    args = ap.parse_args(argv)

    t0 = time.time()
    specs = training_meshes(args.meshes)
    ctxs, pairs, feats = [], [], []
    for i, sp in enumerate(specs):
        m = box_tet10(*sp["grid"], seed=sp["seed"], jitter=0.1, E_sigma=sp["sigma"])
        c = _lib.Context(0)
        c.upload(m)
        feats.append(c.learned_features())
        sp["coarse_nodes"] = coarse_nodes_for(sp["grid"], args.coarse_ratio) if args.coarse_nodes > 0 else 0
        pairs.append(harvest_pairs(c, m, n_pairs=args.pairs, seed=sp["seed"], coarse_nodes=sp["coarse_nodes"]))
        ctxs.append(c)
        if i % 8 == 0:
            log(f"mesh {i}: {sp['grid']} sigma={sp['sigma']:.2f} dofs={m.n_dofs} ({time.time() - t0:.1f}s)")
    lp = LearnedPreconditioner(seed=7)
    lp.fit_normalisation(feats)
    log(f"{len(ctxs)} meshes, {sum(p[0].shape[0] for p in pairs)} pairs; training {args.steps} steps")
    train(ctxs, pairs, lp, steps=args.steps, lr=args.lr, log=log,
          coarse_nodes=[sp["coarse_nodes"] for sp in specs])
    for c in ctxs:
        c.close()

    # ---- evaluation on held-out seeds at full size -------------------------
    results = []
    for k in range(args.eval):
        seed = 9000 + k
        m = box_tet10(*args.eval_grid, seed=seed, jitter=0.1, E_sigma=0.5)
        c = _lib.Context(0)
        c.upload(m)
        c.set_learned_weights(lp.blob())
        row = {"seed": seed}
        cn = args.coarse_nodes
        for name, method, prec in (("fp64_bj", 0, 0), ("fp64_learned", 0, 1), ("mixed_bj", 1, 0), ("mixed_learned", 1, 1)):
            kw = dict(inner_rtol=1e-3, coarse_nodes=cn) if cn > 0 else dict(
                inner_rtol=1e-2, first_rtol=1e-3 if method == 1 else 0.0)
            _, _, st = c.solve(method=method, op="matfree", op_outer="matfree", precond=prec, rtol=1e-8,
                               check_every=32, raise_on_noconv=False, **kw)
            row[name] = {"iterations": st["iterations"], "ms": st["solve_ms"], "setup_ms": st["setup_ms"],
                         "relres": st["relres"], "converged": st["converged"]}
        results.append(row)
        c.close()
        log(f"eval seed {seed}: " + ", ".join(f"{n} {row[n]['iterations']} it {row[n]['ms']:.0f} ms"
                                             for n in ("fp64_bj", "fp64_learned", "mixed_bj", "mixed_learned")))
    summary = {}
    for n in ("fp64_bj", "fp64_learned", "mixed_bj", "mixed_learned"):
        summary[n] = {"mean_iterations": float(np.mean([r[n]["iterations"] for r in results])) if results else None,
                      "mean_ms": float(np.mean([r[n]["ms"] for r in results])) if results else None}
    log(f"summary: {json.dumps(summary)}")
    lp.save(args.out, meta={"training_meshes": specs, "pairs_per_mesh": args.pairs, "steps": args.steps,
                            "coarse_nodes": args.coarse_nodes,
                            "lr": args.lr, "evaluation": results, "summary": summary,
                            "wall_s": time.time() - t0})
    log(f"saved {args.out} ({time.time() - t0:.1f}s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
