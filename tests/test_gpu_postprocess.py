"""Strain/stress recovery on the GPU against the oracle
(src/problems_elasticity.jl:520-594).

Bars: Gauss-point strain and stress within 1e-12 relative (max norm) of the
oracle; nodal least-squares fits (PCG on the mass operator to 1e-13 against
the oracle's sparse direct solve, same fit rule) within 1e-9 relative; an
affine displacement recovered exactly at the nodes.
"""
import sys

import numpy as np
import pytest

from jfem import mesh as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("gen", [M.box_tet10, M.box_hex8, M.box_tet4])
def test_postprocess_matches_oracle(oracle, gen):
    import jfem
    m = gen(6, 2, 2, seed=7, jitter=0.1, E_sigma=0.4)
    say(f"postprocess {m.etype}")
    ctx = jfem.Context(0)
    ctx.upload(m)
    u, _, _ = ctx.solve(method=0, op="matfree", rtol=1e-10)
    out = ctx.postprocess(u)
    strain, stress = oracle.post_gauss(m, u)
    assert np.abs(out["strain"] - strain).max() <= 1e-12 * np.abs(strain).max()
    assert np.abs(out["stress"] - stress).max() <= 1e-12 * np.abs(stress).max()
    for field in ("strain", "stress"):
        ref = oracle.lsq_fit(m, u, field)
        got = out[field + "_nodal"]
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
    ctx.close()


def test_affine_field_recovered_exactly():
    import jfem
    m = M.box_tet10(5, 2, 2, seed=3, jitter=0.15)
    ctx = jfem.Context(0)
    ctx.upload(m)
    A = np.array([[1.0, 0.3, -0.2], [0.1, -0.5, 0.4], [0.25, 0.05, 0.7]]) * 1e-3
    u = (m.X @ A.T).ravel()
    out = ctx.postprocess(u, gauss=False)
    S = 0.5 * (A + A.T)
    ev = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[1, 2], S[0, 2]])
    assert np.abs(out["strain_nodal"] - ev).max() <= 1e-12 * np.abs(ev).max()
    ctx.close()
