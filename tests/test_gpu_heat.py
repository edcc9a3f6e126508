"""Steady heat conduction on the GPU (jfm_heat_solve) against the oracle.

Bars: temperatures within 1e-8 relative (max norm) of the oracle's sparse
direct solve on a jittered mesh with per-element conductivity and source,
boundary flux and convection; the closed-form bar profiles at the nodes
within 1e-9; the elasticity solve unchanged afterwards.
"""
import sys

import numpy as np
import pytest

from jfem import mesh as M

pytestmark = pytest.mark.gpu
L = 10.0


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("gen,ftype", [(M.box_tet10, "Tri6"), (M.box_hex8, "Quad4"), (M.box_tet4, "Tri3")])
def test_heat_matches_oracle(oracle, gen, ftype):
    import jfem
    m = gen(8, 2, 2, seed=4, jitter=0.1)
    say(f"heat {m.etype}")
    ctx = jfem.Context(0)
    ctx.upload(m)
    rng = np.random.default_rng(1)
    k = np.exp(0.5 * rng.standard_normal(m.n_el)) * 3.0
    q = rng.uniform(0.5, 2.0, m.n_el)
    nf = m.faces.shape[0]
    g, h, te = rng.uniform(0, 1, nf), np.full(nf, 0.7), rng.uniform(10, 20, nf)
    fixed = np.isclose(m.X[:, 0], 0.0)
    Tf = np.full(m.n_nodes, 5.0)
    T, st = ctx.heat_solve(k, q, m.faces, flux=g, h=h, T_ext=te, fixed=fixed, T_fixed=Tf, rtol=1e-12)
    Kh, f = oracle.heat_system(m, k, q=q, faces=m.faces, ftype=ftype, flux=g, h=h, T_ext=te)
    To = oracle.heat_solve_direct(Kh, f, fixed, Tf)
    say(f"  {st}")
    assert st["converged"] == 1
    assert np.abs(T - To).max() <= 1e-8 * np.abs(To).max()
    ctx.close()


def test_heat_closed_forms_and_elasticity_untouched(oracle):
    import jfem
    m = M.box_tet10(10, 2, 2, seed=5, jitter=0.1)
    ctx = jfem.Context(0)
    ctx.upload(m)
    u0, _, s0 = ctx.solve(method=0, op="matfree", rtol=1e-9)
    x = m.X[:, 0]
    left, right = np.isclose(x, 0.0), np.isclose(x, L)
    T, _ = ctx.heat_solve(2.0, q=4.0, fixed=left | right, T_fixed=np.zeros(m.n_nodes), rtol=1e-13)
    assert np.abs(T - 4.0 * x * (L - x) / 4.0).max() <= 1e-9 * 50
    T, _ = ctx.heat_solve(2.0, faces=m.faces, h=0.8, T_ext=30.0, fixed=left, T_fixed=np.full(m.n_nodes, 2.0),
                          rtol=1e-13)
    assert np.abs(T - (2.0 + 28.0 * 0.8 * x / (2.0 + 0.8 * L))).max() <= 1e-9 * 30
    u1, _, s1 = ctx.solve(method=0, op="matfree", rtol=1e-9)
    assert s1["iterations"] == s0["iterations"] and np.array_equal(u0, u1)
    ctx.close()
