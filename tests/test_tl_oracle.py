"""Pins the oracle's finite-strain (total Lagrangian, St. Venant-Kirchhoff)
element, oracle.tl_element, the restatement of Problem{Elasticity} with
finite_strain = true (src/problems_elasticity.jl:203-451).  CPU only.

* u = 0: Km is the linear element stiffness of the C oracle, Kg and f_int
  vanish;
* consistency: Km + Kg is the derivative of f_int (central differences),
  which fails for a wrong sign, a transposed F, a missing HᵀH term or a
  dropped Kg;
* objectivity: a rigid rotation R of the deformed state leaves E and S
  unchanged, rotates f_int node by node and Km by blockdiag(R), keeps Kg;
* a pure rotation carries no stress: f_int = 0, Kg = 0;
* the assembled oracle (tl_assemble) equals the element loop, and its small
  load Newton equals the linear solve.
"""
import numpy as np
import pytest

from jfem.mesh import box_hex8, box_tet10

from test_oracle import affine_tet10


def rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def hex8_element(rng):
    X = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float) * 0.5
    return X + 0.05 * rng.standard_normal(X.shape)


CASES = {"Tet10": affine_tet10, "Hex8": hex8_element}


@pytest.mark.parametrize("etype", list(CASES))
def test_tl_zero_state_is_linear(oracle, etype):
    rng = np.random.default_rng(3)
    Xe = CASES[etype](rng)
    Km, Kg, f = oracle.tl_element(etype, Xe, np.zeros(Xe.shape), 210e3, 0.3)
    K0, _ = oracle.element_stiffness(etype, Xe, 210e3, 0.3)
    assert np.abs(Km - K0).max() <= 1e-12 * np.abs(K0).max()
    assert not np.any(Kg) and not np.any(f)


@pytest.mark.parametrize("etype", list(CASES))
def test_tl_tangent_is_derivative_of_force(oracle, etype):
    rng = np.random.default_rng(7)
    Xe = CASES[etype](rng)
    u = 0.05 * rng.standard_normal(Xe.shape)
    E, nu = 1000.0, 0.3
    Km, Kg, f = oracle.tl_element(etype, Xe, u, E, nu)
    assert np.abs(Km - Km.T).max() <= 1e-12 * np.abs(Km).max()
    assert np.abs(Kg - Kg.T).max() <= 1e-12 * np.abs(Kg).max()
    h = 1e-6
    nd = u.size
    Kfd = np.empty((nd, nd))
    for k in range(nd):
        up, um = u.ravel().copy(), u.ravel().copy()
        up[k] += h
        um[k] -= h
        Kfd[:, k] = (oracle.tl_element(etype, Xe, up, E, nu)[2] - oracle.tl_element(etype, Xe, um, E, nu)[2]) / (2 * h)
    K = Km + Kg
    assert np.abs(Kfd - K).max() <= 1e-6 * np.abs(K).max()
    # without Kg the tangent is not consistent: the check above would see it
    assert np.abs(Kfd - Km).max() > 1e-4 * np.abs(Kg).max()


@pytest.mark.parametrize("etype", list(CASES))
def test_tl_objectivity(oracle, etype):
    rng = np.random.default_rng(5)
    Xe = CASES[etype](rng)
    u1 = 0.04 * rng.standard_normal(Xe.shape)
    R = rot([1.0, 2.0, -0.5], 0.7)
    u2 = (Xe + u1) @ R.T - Xe                 # x2 = R x1
    Km1, Kg1, f1 = oracle.tl_element(etype, Xe, u1, 500.0, 0.25)
    Km2, Kg2, f2 = oracle.tl_element(etype, Xe, u2, 500.0, 0.25)
    npe = Xe.shape[0]
    Rb = np.kron(np.eye(npe), R)
    assert np.abs(f2 - Rb @ f1).max() <= 1e-10 * np.abs(f1).max()
    assert np.abs(Km2 - Rb @ Km1 @ Rb.T).max() <= 1e-10 * np.abs(Km1).max()
    assert np.abs(Kg2 - Kg1).max() <= 1e-10 * np.abs(Kg1).max()
    # a pure rotation carries no stress
    Km3, Kg3, f3 = oracle.tl_element(etype, Xe, Xe @ R.T - Xe, 500.0, 0.25)
    K0, _ = oracle.element_stiffness(etype, Xe, 500.0, 0.25)
    assert np.abs(f3).max() <= 1e-11 * np.abs(K0).max() and np.abs(Kg3).max() <= 1e-11 * np.abs(K0).max()
    assert np.abs(Km3 - Rb @ K0 @ Rb.T).max() <= 1e-10 * np.abs(K0).max()


def test_tl_assemble_matches_element_loop(oracle):
    m = box_hex8(3, 2, 2, seed=1, jitter=0.1, E_sigma=0.3)
    u = 1e-2 * np.random.default_rng(2).standard_normal(m.n_dofs)
    Km, Kg, f = oracle.tl_assemble(m, u)
    K = np.zeros((m.n_dofs, m.n_dofs))
    G = np.zeros_like(K)
    fr = np.zeros(m.n_dofs)
    for e, c in enumerate(m.conn):
        dofs = (3 * c[:, None] + np.arange(3)).ravel()
        a, b, fe = oracle.tl_element("Hex8", m.X[c], u.reshape(-1, 3)[c], m.E[e], m.nu[e])
        K[np.ix_(dofs, dofs)] += a
        G[np.ix_(dofs, dofs)] += b
        fr[dofs] -= fe
    import scipy.sparse as sp
    Kd = sp.csr_matrix((Km[2], Km[1], Km[0]), shape=K.shape).toarray()
    Gd = sp.csr_matrix((Kg[2], Kg[1], Kg[0]), shape=K.shape).toarray()
    assert np.abs(Kd - K).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(Gd - G).max() <= 1e-12 * np.abs(G).max()
    assert np.abs(f - fr).max() <= 1e-12 * np.abs(fr).max()
    # the pattern is the linear one (explicit zeros kept)
    lin = oracle.assemble_csr(m)
    assert np.array_equal(Km[0], lin[0]) and np.array_equal(Km[1], lin[1])


def test_tl_small_load_newton_is_linear(oracle):
    m = box_tet10(6, 2, 2, seed=3, jitter=0.1, traction=(0.0, 0.0, -1e-3))
    u, counts = oracle.tl_newton(m, load_steps=1, rtol=1e-10)
    csr = oracle.assemble_csr(m)
    ul, _, _ = oracle.solve_direct(csr, m.fixed_mask(), np.zeros(m.n_dofs), oracle.load_vector(m))
    assert np.linalg.norm(u - ul) / np.linalg.norm(ul) <= 1e-4     # geometric nonlinearity ~ load
    assert counts[0] >= 2
