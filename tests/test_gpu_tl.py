"""Finite-strain assembly (jfm_assemble_tl) through the C ABI against the
oracle's literal restatement of Problem{Elasticity} with finite_strain = true
(oracle.tl_assemble, src/problems_elasticity.jl:203-451), and the legacy
nonlinear loop run!(::Solver{Nonlinear}) (src/solvers.jl:575-620) driven by
it, against the oracle's Newton with direct solves.

Bars: Km and Kg values within 1e-12 (max-norm, relative to the largest
entry), f within 1e-12, the pattern identical to the linear one, Newton
displacements within 1e-8 relative L2.
"""
import sys

import numpy as np
import pytest

from jfem.mesh import box_hex8, box_tet4, box_tet10

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def jf():
    import jfem
    jfem.load()
    return jfem


MESHES = {
    "tet10": lambda: box_tet10(8, 2, 2, seed=5, jitter=0.1, E_sigma=0.5),
    "hex8": lambda: box_hex8(6, 2, 2, seed=2, jitter=0.1, E_sigma=0.3),
    "tet4": lambda: box_tet4(6, 2, 2, seed=4, jitter=0.1, E_sigma=0.3),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_tl_assembly_vs_oracle(jf, oracle, name):
    m = MESHES[name]()
    rng = np.random.default_rng(9)
    u = 2e-2 * rng.standard_normal(m.n_dofs)
    ctx = jf.Context(0)
    ctx.upload_mesh(m.X, m.conn, m.etype, m.E, m.nu)
    say(f"tl assembly {name}")
    vals, kg, f = ctx.assemble_tl(u, geometric=True)
    rowptr, colind = ctx.csr_pattern()
    Km, Kg, fr = oracle.tl_assemble(m, u)
    assert np.array_equal(rowptr, Km[0]) and np.array_equal(colind, Km[1])
    assert np.array_equal(rowptr, Kg[0]) and np.array_equal(colind, Kg[1])
    assert np.abs(vals - Km[2]).max() <= 1e-12 * np.abs(Km[2]).max()
    assert np.abs(kg - Kg[2]).max() <= 1e-12 * np.abs(Kg[2]).max()
    assert np.abs(f - fr).max() <= 1e-12 * np.abs(fr).max()
    # the CSR operator now holds Km + Kg
    x = rng.standard_normal(m.n_dofs)
    ref = oracle.spmv(Km, x) + oracle.spmv(Kg, x)
    y = ctx.apply(x, op="csr")
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    # without geometric stiffness: Km only, and no Kg output
    vals2, kg2, f2 = ctx.assemble_tl(u, geometric=False)
    assert kg2 is None and np.array_equal(vals2, vals) and np.array_equal(f2, f)
    y2 = ctx.apply(x, op="csr")
    ref2 = oracle.spmv(Km, x)
    assert np.abs(y2 - ref2).max() <= 1e-12 * np.abs(ref2).max()
    # the linear assembly afterwards is the linear K again (and EBE uses linear blocks)
    lin = oracle.assemble_csr(m)
    v3, _ = ctx.assemble(want_f=False)
    assert np.abs(v3 - lin[2]).max() <= 1e-12 * np.abs(lin[2]).max()
    ye = ctx.apply(x, op="ebe")
    rl = oracle.spmv(lin, x)
    assert np.abs(ye - rl).max() <= 1e-12 * np.abs(rl).max()


def test_tl_newton_vs_oracle(jf, oracle):
    """run!(::Solver{Nonlinear}) with the GPU assembly: per iteration assemble
    Km + Kg and f = f_ext - f_int(u), solve the eliminated system, u += du."""
    m = box_tet10(8, 2, 2, seed=1, jitter=0.1, traction=(0.0, 0.0, -40.0))
    uo, counts = oracle.tl_newton(m, load_steps=2, rtol=1e-10)
    ctx = jf.Context(0)
    ctx.upload(m)
    fixed = m.fixed_mask() == 1
    fext = oracle.load_vector(m)
    u = np.zeros(m.n_dofs)
    iters = []
    for k in (1, 2):
        lf = k / 2
        # the load factor scales the uploaded traction
        ctx.upload_bcs(m.fixed_mask(), None, m.faces, m.ftype, lf * m.face_traction())
        ref = np.linalg.norm(lf * fext[~fixed])
        for it in range(31):
            _, _, R = ctx.assemble_tl(u, geometric=True, want_vals=False)
            rn = np.linalg.norm(R[~fixed])
            if rn <= 1e-10 * ref:
                break
            assert it < 30
            du, _, st = ctx.solve(f=R, method=0, op="csr", op_outer="csr", rtol=1e-10, check_every=32,
                                  raise_on_noconv=False)
            u += du
        iters.append(it)
        say(f"load step {k}: {it} Newton iterations (oracle {counts[k - 1]}), |R|/|f| = {rn / ref:.2e}")
    tip = u.reshape(-1, 3)[np.isclose(m.X[:, 0], 10.0)].mean(0)
    say(f"tip displacement {tip}")
    assert abs(tip[2]) > 0.5                                   # a large deflection: the nonlinearity matters
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) <= 1e-8
    assert all(abs(a - b) <= 1 for a, b in zip(iters, counts))


def test_host_mirror_nonlinear_run(jf, oracle):
    """The host mirror's run!(::Solver{Nonlinear}) on a finite-strain
    Problem{Elasticity} with geometric stiffness (Assembly.Kg filled, the
    solve on K = Km + Kg), against the oracle's Newton."""
    m = box_tet10(8, 2, 2, seed=1, jitter=0.1, traction=(0.0, 0.0, -20.0))
    uo, _ = oracle.tl_newton(m, load_steps=1, rtol=1e-10)
    body, bc = jf.Problem.from_mesh(m)
    body.properties.finite_strain = True
    body.properties.geometric_stiffness = True
    solver = jf.NonlinearSolver("nl", body, bc)
    solver.convergence_tolerance = 1e-10 * np.linalg.norm(uo)
    solver.max_iterations = 20
    say("host mirror nonlinear run")
    assert jf.run_nonlinear(solver, 0.0, rtol=1e-10)
    assert body.assembly.Kg is not None and abs(body.assembly.Kg).max() > 0
    say(f"  {solver.iteration} iterations, |du| = {solver.u_norm_change:.2e}")
    assert np.linalg.norm(body.u - uo) / np.linalg.norm(uo) <= 1e-8
