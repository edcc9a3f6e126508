"""Neo-Hookean total-Lagrangian path on the GPU against the oracle.

Bars: the FP64 residual lf f_ext - f_int(u) within 1e-11 relative (max norm)
of the oracle's element-by-element restatement; K_T(u) v within 1e-11 in
FP64 and 2e-5 in FP32; the Newton displacement within 1e-7 relative L2 of
the oracle's Newton with direct tangent solves (both converged to 1e-10),
for every solver variant.
"""
import sys

import numpy as np
import pytest

from jfem import mesh as M

pytestmark = pytest.mark.gpu

Q = 52.5     # end traction giving a tip deflection of about 10% of the span


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def ctx_for(m):
    import jfem
    c = jfem.Context(0)
    c.upload(m)
    return c


def masked(v, fx):
    v = v.copy()
    v[fx.astype(bool)] = 0.0
    return v


@pytest.mark.parametrize("gen,grid", [(M.box_tet10, (6, 2, 2)), (M.box_hex8, (6, 2, 2)), (M.box_tet4, (6, 2, 2))])
def test_residual_and_tangent_match_oracle(oracle, gen, grid):
    m = gen(*grid, seed=5, jitter=0.1, E_sigma=0.5, traction=(0.0, 0.0, -Q))
    say(f"residual/tangent {m.etype} {grid}")
    c = ctx_for(m)
    fx = m.fixed_mask()
    rng = np.random.default_rng(11)
    # a smooth bending-like state plus noise, well inside det F > 0
    u = np.zeros((m.n_nodes, 3))
    u[:, 2] = -0.01 * m.X[:, 0] ** 2
    u = (u + 1e-3 * rng.standard_normal(u.shape)).ravel()
    u[fx.astype(bool)] = 0.0
    f = oracle.load_vector(m)
    fint, csr, d = oracle.nh_assemble(m, u, tangent=True)
    assert d > 0
    for lf in (1.0, 0.4):
        R, rn = c.nl_residual(u, lf)
        Ro = masked(lf * f - fint, fx)
        assert np.abs(R - Ro).max() <= 1e-11 * np.abs(fint).max()
        assert rn == pytest.approx(np.linalg.norm(Ro), rel=1e-9)
    v = rng.standard_normal(m.n_dofs)
    y = c.apply(v, op="tangent")
    yo = masked(oracle.spmv(csr, v), fx)
    assert np.abs(y - yo).max() <= 1e-11 * np.abs(yo).max()
    y32 = c.apply(v, op="tangent", precision=1)
    assert np.abs(y32 - yo).max() <= 2e-5 * np.abs(yo).max()
    c.close()


def test_zero_state_tangent_is_linear_operator():
    m = M.box_tet10(6, 2, 2, seed=6, jitter=0.1, E_sigma=0.5)
    c = ctx_for(m)
    c.nl_residual(np.zeros(m.n_dofs), 1.0)
    v = np.random.default_rng(2).standard_normal(m.n_dofs)
    y = c.apply(v, op="tangent")
    yl = c.apply(v, op="matfree")
    assert np.abs(y - yl).max() <= 1e-12 * np.abs(yl).max()
    c.close()


@pytest.fixture(scope="module")
def newton_case(oracle):
    m = M.box_tet10(10, 2, 2, seed=3, jitter=0.1, E_sigma=0.3, traction=(0.0, 0.0, -Q))
    say("oracle Newton 10x2x2")
    uo, counts = oracle.nh_newton(m, load_steps=5, rtol=1e-10)
    return m, uo, counts


@pytest.mark.parametrize("method,precond", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_newton_matches_oracle(newton_case, method, precond):
    from jfem.learned import DEFAULT_WEIGHTS
    m, uo, counts = newton_case
    c = ctx_for(m)
    if precond == 1:
        with open(DEFAULT_WEIGHTS, "rb") as fh:
            c.set_learned_weights(fh.read())
    say(f"GPU Newton method={method} precond={precond}")
    u, st = c.solve_nonlinear(load_steps=5, rtol=1e-10, method=method, precond=precond, learned_probe=8)
    say(f"  {st}")
    assert st["converged"] == 1
    assert np.linalg.norm(u - uo) <= 1e-7 * np.linalg.norm(uo)
    # a full Newton (oracle) takes about 4 iterations per load step; the
    # inexact one may take a few more, not many
    assert st["newton_iterations"] <= 2 * sum(counts) + 5
    # history: one entry per Newton iterate, final one below rtol
    h = c.history(256)
    assert h[-1][1] <= 1e-10
    # the linear path still works afterwards (block Jacobi rebuilt for K)
    ul, _, sl = c.solve(method=0, op="matfree", rtol=1e-8)
    assert sl["converged"] == 1
    c.close()


def test_small_load_newton_is_linear_solve(oracle):
    m = M.box_tet10(8, 2, 2, seed=4, jitter=0.1, traction=(0.0, 0.0, -0.05))
    c = ctx_for(m)
    u, st = c.solve_nonlinear(load_steps=1, rtol=1e-9)
    ul, _, _ = c.solve(method=0, op="matfree", rtol=1e-10)
    # inexact Newton: each tangent solve stops at the forcing tolerance
    # (0.1 first, then Eisenstat-Walker), so a nearly linear step still takes
    # a few iterates
    assert st["converged"] == 1 and st["newton_iterations"] <= 6
    # geometric and material nonlinearity at strain ~1e-5 stay below 1e-4
    assert np.linalg.norm(u - ul) <= 1e-4 * np.linalg.norm(ul)
    c.close()


def test_inadmissible_state_reports_nan():
    m = M.box_tet10(4, 2, 2, seed=1)
    c = ctx_for(m)
    u = np.zeros((m.n_nodes, 3))
    u[:, 0] = -2.0 * m.X[:, 0]          # x -> -x: det F = -1 everywhere
    _, rn = c.nl_residual(u.ravel(), 1.0)
    assert np.isnan(rn)
    c.close()
