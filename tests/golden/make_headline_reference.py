"""Writes headline_160x16x16_seed1.npz: the oracle's FP64 displacement of the
1,048,707-DOF Tet10 cantilever (160x16x16 hexes, seed 1, +-0.1h jitter,
log-normal E with sigma 0.5) in a compact form the GPU tests compare with.

The full vector is 8.4 MB, so the fixture keeps
  * u at every 16th DOF (65,545 values, all three components in turn),
  * |u|_2, the mean end-face deflection of each component,
  * 16 projections w_k . u with +-1 weights from SplitMix64 (stream 40 + k),
which bound the relative L2 error of a candidate both by sampling and,
through E[(w.e)^2] = |e|^2, over all DOFs.

The answer is computed twice by the oracle, by two different iterations on
the oracle's own CSR K, each to a recursive relative residual of 1e-12:
  1. two-level PCG (oracle.twolevel_pcg, aggregates of 2048 nodes);
  2. block-Jacobi PCG (oracle.pcg_bj, C/OpenMP) started from zero.
Their true FP64 residuals |f - K u| / |f| stop at 5e-9 and 2e-8: for this
cantilever the rounding of f - K u itself (|K||u| is far larger than |f|)
sets that floor, so the true residual cannot show the error.  The difference
of the two answers, stored as `cross_check_rel_l2` (1.8e-12 when this file
was written), does: it bounds the fixture's own error.  Run from the
repository root:
    python tests/golden/make_headline_reference.py [--skip-cross-check]
(about 3 minutes, plus about 10 for the cross-check, on 8 cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "juliafem.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from jfem.mesh import box_tet10, splitmix_uniform  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "headline_160x16x16_seed1.npz")
STRIDE = 16
N_PROJ = 16


def projection_weights(n, k):
    return np.where(splitmix_uniform(1, 40 + k, np.arange(n)) < 0.5, -1.0, 1.0)


def summarise(m, u):
    end = np.flatnonzero(np.isclose(m.X[:, 0], 10.0))
    return dict(
        idx_stride=STRIDE,
        u_sample=u[::STRIDE].copy(),
        norm2=np.linalg.norm(u),
        end_mean=u.reshape(-1, 3)[end].mean(axis=0),
        proj=np.array([projection_weights(u.size, k) @ u for k in range(N_PROJ)]),
    )


def true_relres(csr, m, f, u):
    r = f - O.spmv(csr, u)
    r[m.fixed_mask() == 1] = 0.0
    return float(np.linalg.norm(r) / np.linalg.norm(np.where(m.fixed_mask() == 1, 0.0, f)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-cross-check", action="store_true")
    args = ap.parse_args()
    t0 = time.time()
    m = box_tet10(160, 16, 16, seed=1, jitter=0.1, E_sigma=0.5)
    csr = O.assemble_csr(m)
    f = O.load_vector(m)
    print(f"assembled n={m.n_dofs} nnz={csr[1].size} in {time.time() - t0:.1f} s", flush=True)
    t = time.time()
    u, it = O.twolevel_pcg(m, csr, f, 2048, rtol=1e-12)
    rel = true_relres(csr, m, f, u)
    print(f"two-level PCG: {it} iterations, true relres {rel:.2e}, {time.time() - t:.1f} s", flush=True)
    meta = dict(mesh="box_tet10(160,16,16,seed=1,jitter=0.1,E_sigma=0.5)", n_dofs=m.n_dofs,
                nnz=int(csr[1].size), method="oracle.twolevel_pcg(coarse_nodes=2048, rtol=1e-12)",
                iterations=int(it), true_relres=rel)
    if not args.skip_cross_check:
        t = time.time()
        b = np.where(m.fixed_mask() == 1, 0.0, f)
        u2, it2, rel2, rc = O.pcg_bj(csr, m.fixed_mask(), b, rtol=1e-12, maxit=60000)
        rel2 = true_relres(csr, m, f, u2)
        d = float(np.linalg.norm(u2 - u) / np.linalg.norm(u))
        print(f"block-Jacobi PCG: {it2} iterations (rc {rc}), true relres {rel2:.2e}, "
              f"difference {d:.2e}, {time.time() - t:.1f} s", flush=True)
        meta.update(cross_check="oracle.pcg_bj(rtol=1e-12)", cross_check_iterations=int(it2),
                    cross_check_true_relres=rel2, cross_check_rel_l2=d)
    s = summarise(m, u)
    np.savez(OUT, meta=json.dumps(meta), **s)
    print(f"wrote {OUT}: {json.dumps(meta)}")


if __name__ == "__main__":
    main()
