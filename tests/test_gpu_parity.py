"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: CSR pattern bit-exact; K values, loads and FP64 K.u within 1e-12
relative (max-norm); FP32 operators within 2e-6 relative; FP64 solves within
1e-8 relative L2 of the oracle's direct solve; FP32-inner solves to the FP64
residual they report.
"""
import os
import sys

import numpy as np
import pytest

from jfem import mesh as M

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def jf():
    import jfem
    jfem.load()
    return jfem


def c3d10_mesh():
    X, sets = M.read_abaqus_inp(os.path.join(GOLDEN, "contact_3d_tet10.inp"))
    conn = np.vstack([sets["UPPER"][1], sets["LOWER"][1]])
    n = conn.shape[0]
    rng = np.random.default_rng(3)
    return M.Mesh(X=X, conn=conn, etype="Tet10", E=210e3 * np.exp(0.3 * rng.standard_normal(n)),
                  nu=np.full(n, 0.3), faces=np.zeros((0, 6), np.int64), ftype="Tri6",
                  clamp_nodes=np.zeros(0, np.int64))


MESHES = {
    "tet10_cfg0": lambda: M.box_tet10(40, 4, 4),
    "tet10_jitter": lambda: M.box_tet10(12, 3, 3, seed=5, jitter=0.1, E_sigma=0.5),
    "hex8": lambda: M.box_hex8(10, 3, 2, seed=2, jitter=0.1, E_sigma=0.3),
    "tet4": lambda: M.box_tet4(8, 3, 3, seed=4, jitter=0.1, E_sigma=0.3),
    "c3d10_fixture": c3d10_mesh,
}


@pytest.fixture(scope="module")
def cases(jf, oracle):
    out = {}
    for name, make in MESHES.items():
        m = make()
        say(f"upload+assemble {name}: nodes={m.n_nodes} el={m.n_el}")
        ctx = jf.Context(0)
        ctx.upload(m)
        vals, f = ctx.assemble()
        rowptr, colind = ctx.csr_pattern()
        out[name] = dict(mesh=m, ctx=ctx, vals=vals, f=f, rowptr=rowptr, colind=colind,
                         csr=oracle.assemble_csr(m), f_ref=oracle.load_vector(m))
    return out


@pytest.mark.parametrize("name", list(MESHES))
def test_pattern_bit_exact(cases, name):
    c = cases[name]
    assert np.array_equal(c["rowptr"], c["csr"][0])
    assert np.array_equal(c["colind"], c["csr"][1])
    if name == "tet10_cfg0":
        assert c["colind"].size == 1467369


def test_pattern_one_based(cases):
    c = cases["tet4"]
    rp, ci = c["ctx"].csr_pattern(base=1)
    assert np.array_equal(rp, c["csr"][0] + 1) and np.array_equal(ci, c["csr"][1] + 1)


@pytest.mark.parametrize("name", list(MESHES))
def test_values_and_load(cases, name):
    c = cases[name]
    ref = c["csr"][2]
    scale = np.abs(ref).max()
    assert np.abs(c["vals"] - ref).max() <= 1e-12 * scale
    fr = c["f_ref"]
    assert np.abs(c["f"] - fr).max() <= 1e-14 * max(np.abs(fr).max(), 1.0)


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("op", ["csr", "ebe", "matfree"])
def test_apply_fp64_fp32(cases, oracle, name, op):
    c = cases[name]
    m = c["mesh"]
    rng = np.random.default_rng(11)
    x = rng.standard_normal(m.n_dofs)
    ref = oracle.spmv(c["csr"], x)
    ref[m.fixed_mask() == 1] = 0.0
    say(f"apply {name} {op}")
    y = c["ctx"].apply(x, op=op, precision=0)
    s = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 1e-12 * s
    y32 = c["ctx"].apply(x, op=op, precision=1)
    assert np.abs(y32 - ref).max() <= 2e-6 * s


def test_matfree_matches_two_phase_oracle(cases, oracle):
    c = cases["tet10_jitter"]
    m = c["mesh"]
    x = np.random.default_rng(2).standard_normal(m.n_dofs)
    ref = oracle.matfree_apply(m, x, m.fixed_mask())
    y = c["ctx"].apply(x, op="matfree")
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.fixture(scope="module")
def direct0(cases, oracle):
    c = cases["tet10_cfg0"]
    m = c["mesh"]
    u, la, _ = oracle.solve_direct(c["csr"], m.fixed_mask(), np.zeros(m.n_dofs), c["f_ref"])
    return u, la


@pytest.mark.parametrize("op", ["matfree", "csr", "ebe"])
def test_pcg_fp64_vs_direct(cases, direct0, op):
    c = cases["tet10_cfg0"]
    say(f"pcg fp64 cfg0 op={op}")
    u, la, st = c["ctx"].solve(method=0, op=op, rtol=1e-9, want_la=True)
    ud, lad = direct0
    assert st["converged"] == 1 and st["relres"] <= 1e-9
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
    assert np.allclose(la.reshape(-1, 3).sum(0), [0, 0, -1], atol=1e-8)
    fx = c["mesh"].fixed_mask() == 1
    assert np.abs(la[fx] - lad[fx]).max() <= 1e-6 * np.abs(lad).max()


def test_mixed_ir_vs_direct(cases, direct0):
    c = cases["tet10_cfg0"]
    say("mixed IR cfg0")
    u, _, st = c["ctx"].solve(method=1, op="matfree", op_outer="matfree", rtol=1e-9, inner_rtol=1e-4)
    ud, _ = direct0
    assert st["converged"] == 1 and st["relres"] <= 1e-9 and st["outer"] >= 2
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8


def test_solver_bitwise_deterministic(cases):
    c = cases["tet10_jitter"]
    u1, _, s1 = c["ctx"].solve(method=0, op="matfree", rtol=1e-8)
    u2, _, s2 = c["ctx"].solve(method=0, op="matfree", rtol=1e-8)
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(u1, u2)
    x = np.random.default_rng(0).standard_normal(c["mesh"].n_dofs)
    for op in ("csr", "matfree", "ebe"):
        assert np.array_equal(c["ctx"].apply(x, op=op), c["ctx"].apply(x, op=op))


def test_prescribed_displacement(jf, oracle):
    """Nonzero Dirichlet values: u[B] = g[B] and f[I] - K[I,B] g[B] on the right
    (src/solvers.jl:206-210)."""
    m = M.box_hex8(10, 2, 2, seed=1, jitter=0.1)
    fixed = m.fixed_mask()
    far = np.flatnonzero(np.isclose(m.X[:, 0], 10.0))
    g = np.zeros(m.n_dofs)
    fixed[3 * far] = 1
    g[3 * far] = 0.01
    csr = oracle.assemble_csr(m)
    f = np.zeros(m.n_dofs)
    ud, lad, _ = oracle.solve_direct(csr, fixed, g, f)
    ctx = jf.Context(0)
    ctx.upload_mesh(m.X, m.conn, m.etype, m.E, m.nu)
    ctx.upload_bcs(fixed, g)
    say("prescribed displacement hex8")
    for method in (0, 1):
        u, la, st = ctx.solve(f=f, method=method, rtol=1e-10, want_la=True)
        assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
        assert np.array_equal(u[fixed == 1], g[fixed == 1])
        assert np.abs(la - lad).max() <= 1e-6 * np.abs(lad).max()


def test_host_mirror_run(jf, oracle, direct0):
    """run!(analysis) through the Problem / assemble / solve mirror."""
    m = M.box_tet10(40, 4, 4)
    body, bc = jf.Problem.from_mesh(m)
    solver = jf.LinearSolver("solver", body, bc)
    say("host mirror run cfg0")
    u, la = jf.run(solver, 0.0, rtol=1e-9)
    ud, _ = direct0
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
    csr = oracle.assemble_csr(m)
    K = body.assembly.K
    assert np.array_equal(K.indptr, csr[0]) and np.array_equal(K.indices, csr[1])
    # f = f_ext - f_int(u): with the converged u the free rows are in equilibrium
    jf.assemble(body, 0.0) if body.assembly.isempty() else None
    body.assembly.empty()
    body.u = u
    jf.assemble(body, 0.0)
    free = m.fixed_mask() == 0
    assert np.abs(body.assembly.f[free]).max() <= 1e-8 * np.abs(oracle.load_vector(m)).max()


def test_inverted_element_rejected(jf):
    m = M.box_tet10(2, 1, 1)
    conn = m.conn.copy()
    # mirror element 0: swap corners 2,3 and the mid nodes of the swapped edges
    conn[0] = conn[0, [0, 2, 1, 3, 6, 5, 4, 7, 9, 8]]
    ctx = jf.Context(0)
    with pytest.raises(jf.JfemError, match="ENOTSPD"):
        ctx.upload_mesh(m.X, conn, "Tet10", m.E, m.nu)


def test_bad_arguments(jf):
    m = M.box_tet4(2, 1, 1)
    ctx = jf.Context(0)
    with pytest.raises(jf.JfemError, match="EINVAL"):
        ctx.upload_mesh(m.X, m.conn + 10 ** 6, "Tet4", m.E, m.nu)
    with pytest.raises(jf.JfemError, match="ESTATE"):
        ctx.solve()


def test_config1_size_properties(jf):
    """Full 1.05M-DOF mesh: pattern size, operator agreement, equilibrium,
    rigid-body null space and a converged FP64 solve."""
    m = M.box_tet10(160, 16, 16, seed=1, jitter=0.1, E_sigma=0.5)
    ctx = jf.Context(0)
    say("config1 upload (no BCs)")
    ctx.upload_mesh(m.X, m.conn, m.etype, m.E, m.nu)
    inf = ctx.info()
    assert inf["n_dofs"] == 1048707 and inf["nnz"] == 87026697 and inf["affine"] == 1
    rng = np.random.default_rng(5)
    x = rng.standard_normal(m.n_dofs)
    y_mf = ctx.apply(x, op="matfree")
    y_csr = ctx.apply(x, op="csr")
    s = np.abs(y_csr).max()
    assert np.abs(y_mf - y_csr).max() <= 1e-12 * s
    assert np.abs(y_mf.reshape(-1, 3).sum(0)).max() <= 1e-10 * s * np.sqrt(m.n_nodes)
    for axis in np.eye(3):
        r = np.cross(axis, m.X).ravel()
        assert np.abs(ctx.apply(r, op="matfree")).max() <= 1e-9 * s * np.abs(r).max()
    say("config1 solve fp64 matfree")
    ctx.upload_bcs(m.fixed_mask(), None, m.faces, m.ftype, m.face_traction())
    u, _, st = ctx.solve(method=0, op="matfree", rtol=1e-8)
    assert st["converged"] == 1 and st["relres"] <= 1e-8
    say(f"config1 pcg iterations={st['iterations']} ms={st['solve_ms']:.1f}")
    u2, _, st2 = ctx.solve(method=1, op="matfree", op_outer="matfree", rtol=1e-8)
    assert st2["converged"] == 1
    assert np.linalg.norm(u2 - u) / np.linalg.norm(u) <= 1e-6
    say(f"config1 mixed iterations={st2['iterations']} outer={st2['outer']} ms={st2['solve_ms']:.1f}")
