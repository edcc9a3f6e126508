"""Parity of the headline workload with the oracle, through the C ABI.

* 80x8x8 jittered Tet10 cantilever with log-normal E (sigma 0.5), 139,587
  DOFs: every solve method of the benchmark, including the headline options,
  against the oracle's FP64 banded-Cholesky direct solve (one refinement
  step), within 1e-8 relative L2.
* 160x16x16 (1,048,707 DOFs, the benchmark mesh): the CSR pattern bit-exact
  and the K values within 1e-12 of the oracle's assembly; the headline
  solve and the FP64 two-level solve against the oracle's FP64 answer kept
  in tests/golden/headline_160x16x16_seed1.npz (written by
  tests/golden/make_headline_reference.py), within 1e-8 relative L2.
"""
import json
import os
import sys

import numpy as np
import pytest

from jfem.mesh import box_tet10, splitmix_uniform

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# the benchmark's methods: (jfm_solve method, preconditioner, options)
HEADLINE = dict(method=1, precond=1, learned_probe=16, inner_rtol=2e-3, coarse_nodes=1536, check_every=16)
METHODS = {
    "mixed-learned-2l": HEADLINE,
    "mixed-bj-2l": dict(method=1, precond=0, inner_rtol=2e-3, coarse_nodes=1536, check_every=16),
    "fp64-bj-2l": dict(method=0, precond=0, coarse_nodes=1536, check_every=16),
    "fp64-bj": dict(method=0, precond=0),
}


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def default_weights():
    from jfem.learned import DEFAULT_WEIGHTS
    with open(DEFAULT_WEIGHTS, "rb") as fh:
        return fh.read()


@pytest.fixture(scope="module")
def jf():
    import jfem
    jfem.load()
    return jfem


# ------------------------------------------------------------ 80 x 8 x 8 --
@pytest.fixture(scope="module")
def mid(jf, oracle):
    m = box_tet10(80, 8, 8, seed=7, jitter=0.1, E_sigma=0.5)
    say(f"80x8x8 oracle assembly and direct solve: n={m.n_dofs}")
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    ud, lad, bw = oracle.solve_direct_lapack(csr, m.fixed_mask(), np.zeros(m.n_dofs), f)
    ctx = jf.Context(0)
    ctx.upload(m)
    ctx.set_learned_weights(default_weights())
    return dict(m=m, csr=csr, f=f, ud=ud, lad=lad, ctx=ctx)


def test_mid_pattern_and_values(mid):
    vals, f = mid["ctx"].assemble()
    rowptr, colind = mid["ctx"].csr_pattern()
    csr = mid["csr"]
    assert np.array_equal(rowptr, csr[0]) and np.array_equal(colind, csr[1])
    assert np.abs(vals - csr[2]).max() <= 1e-12 * np.abs(csr[2]).max()
    assert np.abs(f - mid["f"]).max() <= 1e-14 * np.abs(mid["f"]).max()


@pytest.mark.parametrize("name", list(METHODS))
def test_mid_solve_vs_direct(mid, name):
    o = METHODS[name]
    say(f"80x8x8 solve {name}")
    u, la, st = mid["ctx"].solve(method=o["method"], precond=o["precond"], rtol=1e-8,
                                 inner_rtol=o.get("inner_rtol", 1e-4), learned_probe=o.get("learned_probe", 0),
                                 coarse_nodes=o.get("coarse_nodes", 0), check_every=o.get("check_every", 32),
                                 want_la=True)
    ud = mid["ud"]
    err = np.linalg.norm(u - ud) / np.linalg.norm(ud)
    say(f"  iterations={st['iterations']} relres={st['relres']:.2e} err_vs_direct={err:.2e} "
        f"fallback={st['fallback']}")
    assert st["converged"] == 1 and st["relres"] <= 1e-8
    assert err <= 1e-8
    fx = mid["m"].fixed_mask() == 1
    assert np.abs(la[fx] - mid["lad"][fx]).max() <= 1e-6 * np.abs(mid["lad"]).max()


# ------------------------------------------------------- 160 x 16 x 16 --
@pytest.fixture(scope="module")
def big(jf):
    m = box_tet10(160, 16, 16, seed=1, jitter=0.1, E_sigma=0.5)
    ctx = jf.Context(0)
    say("160x16x16 upload")
    ctx.upload(m)
    ctx.set_learned_weights(default_weights())
    return dict(m=m, ctx=ctx)


def test_config1_pattern_and_values_vs_oracle(big, oracle):
    m, ctx = big["m"], big["ctx"]
    say("160x16x16 GPU assembly")
    vals, f = ctx.assemble()
    rowptr, colind = ctx.csr_pattern()
    say("160x16x16 oracle assembly")
    csr = oracle.assemble_csr(m)
    assert rowptr.size == m.n_dofs + 1 and colind.size == 87026697
    assert np.array_equal(rowptr, csr[0])
    assert np.array_equal(colind, csr[1])
    del rowptr, colind
    assert np.abs(vals - csr[2]).max() <= 1e-12 * np.abs(csr[2]).max()
    fr = oracle.load_vector(m)
    assert np.abs(f - fr).max() <= 1e-14 * np.abs(fr).max()


def _fixture():
    z = np.load(os.path.join(GOLDEN, "headline_160x16x16_seed1.npz"))
    return {k: z[k] for k in z.files}


def _compare(m, u, ref):
    """(sampled relative L2 error, projection estimate of the full relative
    L2 error, |u| relative error, end-face deflection relative error)."""
    s = int(ref["idx_stride"])
    us = u[::s]
    e_sample = np.linalg.norm(us - ref["u_sample"]) / np.linalg.norm(ref["u_sample"])
    proj = np.array([np.where(splitmix_uniform(1, 40 + k, np.arange(u.size)) < 0.5, -1.0, 1.0) @ u
                     for k in range(ref["proj"].size)])
    e_proj = np.sqrt(np.mean((proj - ref["proj"]) ** 2)) / float(ref["norm2"])
    e_norm = abs(np.linalg.norm(u) - float(ref["norm2"])) / float(ref["norm2"])
    end = np.flatnonzero(np.isclose(m.X[:, 0], 10.0))
    em = u.reshape(-1, 3)[end].mean(axis=0)
    e_end = np.abs(em - ref["end_mean"]).max() / np.abs(ref["end_mean"]).max()
    return e_sample, e_proj, e_norm, e_end


@pytest.mark.parametrize("name", ["mixed-learned-2l", "fp64-bj-2l"])
def test_config1_solve_vs_oracle_fixture(big, name):
    ref = _fixture()
    meta = json.loads(str(ref["meta"]))
    assert meta["n_dofs"] == big["m"].n_dofs
    o = METHODS[name]
    say(f"160x16x16 solve {name}")
    u, _, st = big["ctx"].solve(method=o["method"], precond=o["precond"], rtol=1e-8,
                                inner_rtol=o.get("inner_rtol", 1e-4), learned_probe=o.get("learned_probe", 0),
                                coarse_nodes=o.get("coarse_nodes", 0), check_every=o.get("check_every", 32))
    e = _compare(big["m"], u, ref)
    say(f"  iterations={st['iterations']} relres={st['relres']:.2e} err sample={e[0]:.2e} "
        f"projections={e[1]:.2e} norm={e[2]:.2e} end={e[3]:.2e}")
    assert st["converged"] == 1 and st["relres"] <= 1e-8
    assert max(e) <= 1e-8
