"""Pins the CPU oracle against closed forms, invariants, the reference's own
known answers and an independent library solver.  CPU only.

Every check is chosen so that a plausible slip in the restatement (a wrong
shear factor, a transposed Jacobian, a dropped weight, a mis-ordered mid
node) fails at least one of them.
"""
import os

import numpy as np
import pytest

from jfem.mesh import TET10_EDGES, box_hex8, box_tet10, read_abaqus_inp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# Reference Tet10 nodes (src/basis/lagrange_generated.jl:261-266)
REF_TET10 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0],
                      [0, .5, 0], [0, 0, .5], [.5, 0, .5], [0, .5, .5]], dtype=float)
REF_HEX8 = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
REF_TRI6 = np.array([[0, 0], [1, 0], [0, 1], [.5, 0], [.5, .5], [0, .5]], dtype=float)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def rigid_modes(X):
    """6 rigid-body displacement fields (3 translations, 3 small rotations)."""
    n = X.shape[0]
    modes = []
    for c in range(3):
        m = np.zeros((n, 3))
        m[:, c] = 1
        modes.append(m.ravel())
    for axis in np.eye(3):
        modes.append(np.cross(axis, X).ravel())
    return np.array(modes)


def affine_tet10(rng):
    corners = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + 0.15 * rng.standard_normal((4, 3))
    X = np.zeros((10, 3))
    X[:4] = corners
    for k, (a, b) in enumerate(TET10_EDGES):
        X[4 + k] = 0.5 * (corners[a] + corners[b])
    return X


def tet_volume(c):
    return np.linalg.det(np.stack([c[1] - c[0], c[2] - c[0], c[3] - c[0]])) / 6


# ------------------------------------------------------------- quadrature --

def test_tet10_rule_moments(oracle):
    w, xi = oracle.volume_rule("Tet10")
    assert w.size == 4 and np.allclose(w, 1 / 24)
    x, y, z = xi.T
    exact = {"1": 1 / 6, "x": 1 / 24, "xx": 1 / 60, "xy": 1 / 120, "yz": 1 / 120, "zz": 1 / 60}
    got = {"1": w.sum(), "x": w @ x, "xx": w @ x ** 2, "xy": w @ (x * y), "yz": w @ (y * z), "zz": w @ z ** 2}
    for k in exact:
        assert got[k] == pytest.approx(exact[k], rel=1e-14, abs=0), k


def test_hex8_rule_moments(oracle):
    w, xi = oracle.volume_rule("Hex8")
    x, y, z = xi.T
    assert w.sum() == pytest.approx(8.0, rel=1e-15)
    assert w @ (x ** 2 * y ** 2 * z ** 2) == pytest.approx((2 / 3) ** 3, rel=1e-14)
    assert abs(w @ (x ** 3 * y)) < 1e-15


def test_surface_rules(oracle):
    w, xi = oracle.surface_rule("Tri6")
    assert w.sum() == pytest.approx(0.5) and w @ (xi[:, 0] ** 2) == pytest.approx(1 / 12, rel=1e-14)
    w, xi = oracle.surface_rule("Quad4")
    assert w.sum() == pytest.approx(4.0) and w @ (xi[:, 0] ** 2 * xi[:, 1] ** 2) == pytest.approx(4 / 9, rel=1e-14)


# ------------------------------------------------------------------ basis --

@pytest.mark.parametrize("etype,ref,surface", [("Tet10", REF_TET10, False), ("Hex8", REF_HEX8, False),
                                               ("Tri6", REF_TRI6, True)])
def test_basis_kronecker_partition_fd(oracle, etype, ref, surface):
    n = ref.shape[0]
    for j in range(n):
        N, _ = oracle.basis(etype, ref[j], surface)
        assert np.allclose(N, np.eye(n)[j], atol=1e-15)
    rng = np.random.default_rng(1)
    for _ in range(5):
        xi = rng.uniform(0.05, 0.3, ref.shape[1])
        N, dN = oracle.basis(etype, xi, surface)
        assert N.sum() == pytest.approx(1.0, abs=1e-14)
        assert np.allclose(dN.sum(axis=0), 0, atol=1e-14)
        h = 1e-6
        for a in range(ref.shape[1]):
            e = np.zeros(ref.shape[1])
            e[a] = h
            fd = (oracle.basis(etype, xi + e, surface)[0] - oracle.basis(etype, xi - e, surface)[0]) / (2 * h)
            assert np.allclose(fd, dN[:, a], atol=1e-8)


# ------------------------------------------------------ element stiffness --

def check_element(oracle, etype, X, E=210e3, nu=0.3, affine=True):
    K, det = oracle.element_stiffness(etype, X, E, nu)
    assert det > 0
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-13 * scale
    R = rigid_modes(X)
    assert np.abs(K @ R.T).max() <= 1e-11 * scale * np.abs(R).max()
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert (ev[:6] < 1e-9 * ev[-1]).all() and (ev[6:] > 1e-9 * ev[-1]).all()
    if affine:
        # u = A X gives the constant strain eps = sym(A); u'Ku = V eps:sigma.
        rng = np.random.default_rng(3)
        A = rng.standard_normal((3, 3))
        u = (X @ A.T).ravel()
        eps = 0.5 * (A + A.T)
        la, mu = lame(E, nu)
        sig = la * np.trace(eps) * np.eye(3) + 2 * mu * eps
        vol = tet_volume(X[:4]) if etype.startswith("Tet") else abs(np.linalg.det(
            np.stack([X[1] - X[0], X[3] - X[0], X[4] - X[0]])))
        assert u @ K @ u == pytest.approx(vol * np.sum(eps * sig), rel=1e-12)
        fe = oracle.element_internal_force(etype, X, u, E, nu)
        assert np.allclose(fe, K @ u, rtol=0, atol=1e-12 * np.abs(K @ u).max())
    return K


def test_tet10_stiffness_invariants(oracle):
    rng = np.random.default_rng(0)
    for _ in range(4):
        check_element(oracle, "Tet10", affine_tet10(rng), E=rng.uniform(1e3, 1e6), nu=rng.uniform(0.1, 0.45))


def test_tet10_curved_stiffness_invariants(oracle):
    rng = np.random.default_rng(5)
    X = affine_tet10(rng)
    X[4:] += 0.03 * rng.standard_normal((6, 3))     # curved edges: J varies over the element
    check_element(oracle, "Tet10", X, affine=False)


def test_hex8_stiffness_invariants(oracle):
    rng = np.random.default_rng(2)
    M = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    X = (0.5 * (REF_HEX8 + 1)) @ M.T + rng.standard_normal(3)   # parallelepiped (affine)
    check_element(oracle, "Hex8", X)
    X2 = X + 0.05 * rng.standard_normal((8, 3))
    check_element(oracle, "Hex8", X2, affine=False)


def test_tet4_known_answer(oracle):
    """Unit Tet4, E=210e3, nu=0.3, u_2x = 0.01 (demos/nodal_assembly_cpu.jl:294-316);
    the linear-elastic residual norm in closed form is 779.1428684570344."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    K, _ = oracle.element_stiffness("Tet4", X, 210e3, 0.3)
    u = np.zeros(12)
    u[3] = 0.01
    r = K @ u
    assert np.linalg.norm(r) == pytest.approx(779.1428684570344, rel=1e-14)
    assert np.allclose(r.reshape(4, 3).sum(axis=0), 0, atol=1e-10)


def test_tet4_plastic_known_answer(oracle):
    """The reference demo's own printed answer, 727.208..., for the same element
    with von Mises radial return (sigma_y = 250, demos/nodal_assembly_cpu.jl:78-116).
    The oracle supplies grad N and detJ (through K of unit probes); the return
    map is the demo's closed form."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    E, nu, sy = 210e3, 0.3, 250.0
    la, mu = lame(E, nu)
    vol = 1 / 6
    # grad N_i = invJ . dN_i with J = sum dN_i (x) X_i (src/basis/math.jl:185-212)
    _, dN = oracle.basis("Tet4", [0.25, 0.25, 0.25])
    J = dN.T @ X
    G = dN @ np.linalg.inv(J).T
    eps = np.zeros((3, 3))
    eps[0, 0] = 0.01 * G[1, 0]
    eps[0, 1] = eps[1, 0] = 0.5 * 0.01 * G[1, 1]
    eps[0, 2] = eps[2, 0] = 0.5 * 0.01 * G[1, 2]
    sig = la * np.trace(eps) * np.eye(3) + 2 * mu * eps
    dev = sig - np.trace(sig) / 3 * np.eye(3)
    seq = np.sqrt(1.5 * np.sum(dev * dev))
    f = seq - sy
    if f > 0:
        sig = sig - 2 * mu * (f / (3 * mu)) * dev / seq
    r = np.concatenate([vol * G[i] @ sig for i in range(4)])
    assert np.linalg.norm(r) == pytest.approx(727.2081516082287, rel=1e-12)
    # and the elastic branch of the same path reproduces the oracle's K u
    K, _ = oracle.element_stiffness("Tet4", X, E, nu)
    u = np.zeros(12)
    u[3] = 0.01
    sig_e = la * np.trace(eps) * np.eye(3) + 2 * mu * eps
    r_e = np.concatenate([vol * G[i] @ sig_e for i in range(4)])
    assert np.allclose(r_e, K @ u, rtol=1e-13, atol=1e-9)


# ---------------------------------------------------------- surface loads --

def test_surface_consistent_loads(oracle):
    rng = np.random.default_rng(4)
    t = np.array([0.3, -0.2, -1.0])
    P = rng.standard_normal((3, 3))
    A = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    Xf = np.vstack([P, 0.5 * (P[0] + P[1]), 0.5 * (P[1] + P[2]), 0.5 * (P[2] + P[0])])
    fe = oracle.surface_load("Tri6", Xf, t).reshape(6, 3)
    assert np.allclose(fe[:3], 0, atol=1e-15)
    assert np.allclose(fe[3:], A / 3 * t, rtol=1e-13)
    fe = oracle.surface_load("Tri3", P, t).reshape(3, 3)
    assert np.allclose(fe, A / 3 * t, rtol=1e-13)
    Q = np.vstack([P[0], P[1], P[1] + P[2] - P[0], P[2]])         # parallelogram
    Aq = np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    fe = oracle.surface_load("Quad4", Q, t).reshape(4, 3)
    assert np.allclose(fe, Aq / 4 * t, rtol=1e-13)


# ---------------------------------------------------------- global system --

@pytest.fixture(scope="module")
def cfg0(oracle):
    m = box_tet10(40, 4, 4)
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    return m, csr, f


def brute_pattern(conn, n_nodes):
    adj = [set() for _ in range(n_nodes)]
    for el in conn:
        for a in el:
            adj[a].update(el.tolist())
    rowptr = [0]
    cols = []
    for a in range(n_nodes):
        nb = sorted(adj[a])
        blk = [3 * b + d for b in nb for d in range(3)]
        for c in range(3):
            cols.extend(blk)
            rowptr.append(len(cols))
    return np.array(rowptr), np.array(cols)


def test_config0_pattern_bruteforce(cfg0):
    m, csr, _ = cfg0
    assert m.n_nodes == 6561 and m.n_dofs == 19683 and m.n_el == 3840
    rp, ci = brute_pattern(m.conn, m.n_nodes)
    assert csr[1].size == 1467369
    assert np.array_equal(csr[0], rp) and np.array_equal(csr[1], ci)


def test_config0_load_total(cfg0):
    m, _, f = cfg0
    assert np.allclose(f.reshape(-1, 3).sum(axis=0), [0, 0, -1], atol=1e-14)


def test_equilibrium_and_symmetry(cfg0, oracle):
    m, csr, _ = cfg0
    rng = np.random.default_rng(7)
    u = rng.standard_normal(m.n_dofs)
    y = oracle.spmv(csr, u)
    assert np.allclose(y.reshape(-1, 3).sum(axis=0), 0, atol=1e-9 * np.abs(y).max())
    import scipy.sparse as sp
    K = sp.csr_matrix((csr[2], csr[1], csr[0]))
    assert abs(K - K.T).max() <= 1e-12 * abs(K).max()
    for r in rigid_modes(m.X):
        assert np.abs(oracle.spmv(csr, r)).max() <= 1e-9 * abs(K).max()


def test_direct_solve_superlu_and_beam(cfg0, oracle):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    m, csr, f = cfg0
    fx = m.fixed_mask()
    u, la, bw = oracle.solve_direct(csr, fx, np.zeros(m.n_dofs), f)
    free = np.flatnonzero(fx == 0)
    K = sp.csr_matrix((csr[2], csr[1], csr[0]))
    r = f[free] - K[free] @ u
    assert np.linalg.norm(r) / np.linalg.norm(f) < 5e-9
    us = sla.spsolve(K[free][:, free].tocsc(), f[free])
    assert np.linalg.norm(us - u[free]) / np.linalg.norm(us) < 1e-8
    assert np.all(u[fx == 1] == 0)
    # reactions balance the applied load
    assert np.allclose(la.reshape(-1, 3).sum(axis=0), [0, 0, -1], atol=1e-9)
    # Euler-Bernoulli tip deflection F L^3 / (3 E I) with I = 1/12
    tip = u.reshape(-1, 3)[-1, 2]
    assert abs(-tip / (1 * 10 ** 3 / (3 * 210e3 / 12)) - 1) < 0.02


def test_two_phase_matches_csr(oracle):
    m = box_tet10(12, 3, 3, seed=11, jitter=0.1, E_sigma=0.5)
    csr = oracle.assemble_csr(m)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(m.n_dofs)
    fx = m.fixed_mask()
    y1 = oracle.matfree_apply(m, u, fx)
    y2 = oracle.spmv(csr, u)
    y2[fx == 1] = 0
    assert np.abs(y1 - y2).max() <= 1e-12 * np.abs(y2).max()


def test_pcg_matches_direct(oracle):
    m = box_tet10(12, 3, 3, seed=12, jitter=0.1, E_sigma=0.5)
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    fx = m.fixed_mask()
    u, _, _ = oracle.solve_direct(csr, fx, np.zeros(m.n_dofs), f)
    x, it, rel, rc = oracle.pcg_bj(csr, fx, f, rtol=1e-11)
    assert rc == 0 and rel <= 1e-11
    assert np.linalg.norm(x - u) / np.linalg.norm(u) < 1e-8


def test_hex8_mesh_system(oracle):
    m = box_hex8(10, 2, 2, seed=3, jitter=0.1)
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    assert np.allclose(f.reshape(-1, 3).sum(axis=0), [0, 0, -1], atol=1e-14)
    rp, ci = brute_pattern(m.conn, m.n_nodes)
    assert np.array_equal(csr[0], rp) and np.array_equal(csr[1], ci)


def test_reference_c3d10_fixture(oracle):
    """The reference's own irregular C3D10 mesh (test/test_problems_contact_3d/tet10.inp)."""
    X, sets = read_abaqus_inp(os.path.join(GOLDEN, "contact_3d_tet10.inp"))
    assert X.shape == (607, 3)
    assert sets["UPPER"][1].shape == (177, 10) and sets["LOWER"][1].shape == (92, 10)
    from jfem.mesh import Mesh
    conn = np.vstack([sets["UPPER"][1], sets["LOWER"][1]])
    n_el = conn.shape[0]
    m = Mesh(X=X, conn=conn, etype="Tet10", E=np.full(n_el, 210e3), nu=np.full(n_el, 0.3),
             faces=np.zeros((0, 6), np.int64), ftype="Tri6", clamp_nodes=np.zeros(0, np.int64))
    csr = oracle.assemble_csr(m)
    import scipy.sparse as sp
    K = sp.csr_matrix((csr[2], csr[1], csr[0]), shape=(m.n_dofs, m.n_dofs))
    assert abs(K - K.T).max() <= 1e-12 * abs(K).max()
    used = np.unique(conn)
    Xu = np.zeros_like(X)
    Xu[used] = X[used]
    for r in rigid_modes(Xu):
        assert np.abs(K @ r).max() <= 1e-9 * abs(K).max() * np.abs(r).max()
