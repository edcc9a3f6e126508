"""Pins the oracle's neo-Hookean total-Lagrangian restatement.  CPU only.

The reference obtains S and D by automatic differentiation of psi
(src/materials/neo_hookean.jl:205-231); the oracle writes them in closed form,
so they are checked against central differences of psi and of S.  The TL
element (src/problems_elasticity.jl:246-289, 369-409) is checked against
differences of its own force, against the already-pinned linear element at
u = 0, for objectivity (rigid translation and finite rotation give zero
force), and by the patch test.  A transposed F in BL(F), a missing geometric
stiffness or a sign slip in ln J fails at least one of them.
"""
import numpy as np
import pytest

from jfem.mesh import box_tet10

from test_oracle import REF_HEX8, REF_TET10, lame

MU, LA = 80.77e3, 121.15e3


def rand_E(rng, amp=0.3):
    """Green-Lagrange strain of a random admissible F."""
    F = np.eye(3) + amp * rng.standard_normal((3, 3))
    if np.linalg.det(F) < 0:
        F[:, 0] *= -1
    return 0.5 * (F.T @ F - np.eye(3))


def sym_dir(i, j):
    d = np.zeros((3, 3))
    d[i, j] = d[j, i] = 1.0
    return d


VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]


def test_stress_is_derivative_of_energy(oracle):
    rng = np.random.default_rng(1)
    for _ in range(5):
        E = rand_E(rng)
        _, S, _ = oracle.nh_material(E, MU, LA)
        h = 1e-6
        for i, j in VOIGT:
            d = sym_dir(i, j)
            dpsi = (oracle.nh_material(E + h * d, MU, LA)[0] - oracle.nh_material(E - h * d, MU, LA)[0]) / (2 * h)
            # dpsi = S:dE; a symmetric off-diagonal direction counts S_ij twice
            want = 2 * S[i, j] if i != j else S[i, i]
            assert dpsi == pytest.approx(want, rel=1e-6, abs=1e-4)


def test_tangent_is_derivative_of_stress(oracle):
    rng = np.random.default_rng(2)
    for _ in range(5):
        E = rand_E(rng)
        _, _, D = oracle.nh_material(E, MU, LA)
        h = 1e-7
        for J, (k, l) in enumerate(VOIGT):
            # unit engineering strain in component J (shear: E_kl = E_lk = 1/2)
            dE = sym_dir(k, k) if k == l else 0.5 * sym_dir(k, l)
            dS = (oracle.nh_material(E + h * dE, MU, LA)[1] - oracle.nh_material(E - h * dE, MU, LA)[1]) / (2 * h)
            for I, (i, j) in enumerate(VOIGT):
                assert dS[i, j] == pytest.approx(D[I, J], rel=1e-5, abs=1e-2)


def test_small_strain_limit_is_linear_D(oracle):
    la, mu = lame(210e3, 0.3)
    psi, S, D = oracle.nh_material(np.zeros((3, 3)), mu, la)
    assert psi == 0.0 and np.abs(S).max() == 0.0
    Dl = np.zeros((6, 6))
    Dl[:3, :3] = la
    Dl[np.arange(3), np.arange(3)] += 2 * mu
    Dl[np.arange(3, 6), np.arange(3, 6)] = mu
    assert np.allclose(D, Dl, rtol=1e-14, atol=1e-9)
    # the material parameters of the nonlinear workload are E = 210e3, nu = 0.3
    assert la == pytest.approx(121.15e3, rel=1e-4) and mu == pytest.approx(80.77e3, rel=1e-4)


def test_inadmissible_state_is_nan(oracle):
    E = np.diag([0.0, 0.0, -0.75])     # C = diag(1, 1, -0.5)
    assert np.isnan(oracle.nh_material(E, MU, LA)[0])


@pytest.mark.parametrize("etype,Xref", [("Tet10", REF_TET10), ("Hex8", REF_HEX8)])
def test_element_tangent_matches_force_differences(oracle, etype, Xref):
    rng = np.random.default_rng(3)
    Xe = (Xref * [1.3, 0.9, 1.1] + 0.03 * rng.standard_normal(Xref.shape)).ravel()
    ue = 0.05 * rng.standard_normal(Xe.size)
    f0, K, d = oracle.nh_element(etype, Xe, ue, 210e3, 0.3)
    assert d > 0
    h = 1e-7
    Kfd = np.empty_like(K)
    for c in range(Xe.size):
        e = np.zeros(Xe.size)
        e[c] = h
        Kfd[:, c] = (oracle.nh_element(etype, Xe, ue + e, 210e3, 0.3, False)[0]
                     - oracle.nh_element(etype, Xe, ue - e, 210e3, 0.3, False)[0]) / (2 * h)
    assert np.abs(K - Kfd).max() <= 1e-6 * np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-10 * np.abs(K).max()


@pytest.mark.parametrize("etype,Xref", [("Tet10", REF_TET10), ("Hex8", REF_HEX8)])
def test_zero_displacement_is_linear_element(oracle, etype, Xref):
    rng = np.random.default_rng(4)
    Xe = (Xref + 0.05 * rng.standard_normal(Xref.shape)).ravel()
    f, K, d = oracle.nh_element(etype, Xe, np.zeros(Xe.size), 210e3, 0.3)
    Kl, _ = oracle.element_stiffness(etype, Xe, 210e3, 0.3)
    assert d == 1.0 and np.abs(f).max() == 0.0
    assert np.abs(K - Kl).max() <= 1e-12 * np.abs(Kl).max()


def test_objectivity(oracle):
    rng = np.random.default_rng(5)
    X = (REF_TET10 + 0.05 * rng.standard_normal(REF_TET10.shape))
    # rigid translation
    f, _, _ = oracle.nh_element("Tet10", X.ravel(), np.tile([0.3, -0.2, 0.7], 10), 210e3, 0.3, False)
    assert np.abs(f).max() <= 1e-9
    # finite rotation (60 degrees about a skew axis): u = (R - I) X
    ax = np.array([1.0, 2.0, -0.5])
    ax /= np.linalg.norm(ax)
    th = np.pi / 3
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx
    u = (X @ R.T - X).ravel()
    f, _, d = oracle.nh_element("Tet10", X.ravel(), u, 210e3, 0.3, False)
    assert d == pytest.approx(1.0, abs=1e-12)
    assert np.abs(f).max() <= 1e-8 * 210e3


def test_patch_test_homogeneous_deformation(oracle):
    m = box_tet10(3, 2, 2, seed=8, jitter=0.1)
    Fh = np.array([[1.08, 0.05, -0.02], [0.03, 0.96, 0.04], [-0.01, 0.02, 1.05]])
    u = (m.X @ (Fh - np.eye(3)).T).ravel()
    fint, _, d = oracle.nh_assemble(m, u, tangent=False)
    assert d > 0
    xs = m.X
    tol = 1e-9 * 210e3
    interior = np.all((xs > xs.min(axis=0) + 1e-9) & (xs < xs.max(axis=0) - 1e-9), axis=1)
    assert interior.sum() > 0
    assert np.abs(fint.reshape(-1, 3)[interior]).max() <= tol
    # total force on the body vanishes (sum of all nodal forces)
    assert np.abs(fint.reshape(-1, 3).sum(axis=0)).max() <= tol


def test_assembled_tangent_is_pattern_of_linear(oracle):
    m = box_tet10(3, 2, 2, seed=9, jitter=0.1)
    u = 1e-3 * np.random.default_rng(1).standard_normal(m.n_dofs)
    _, csr, _ = oracle.nh_assemble(m, u, tangent=True)
    lin = oracle.assemble_csr(m)
    assert np.array_equal(csr[0], lin[0]) and np.array_equal(csr[1], lin[1])
    _, csr0, _ = oracle.nh_assemble(m, np.zeros(m.n_dofs), tangent=True)
    assert np.abs(csr0[2] - lin[2]).max() <= 1e-12 * np.abs(lin[2]).max()


def test_small_strain_stress_keeps_relative_precision(oracle):
    """S at strain 1e-9 equals D:E to rounding (no cancellation in I - C^-1)."""
    la, mu = lame(210e3, 0.3)
    E = 1e-9 * np.array([[1.0, 0.3, -0.2], [0.3, -0.5, 0.1], [-0.2, 0.1, 0.7]])
    _, S, D = oracle.nh_material(E, mu, la)
    lin = la * np.trace(E) * np.eye(3) + 2 * mu * E
    assert np.abs(S - lin).max() <= 1e-7 * np.abs(lin).max()


def test_oracle_newton_small_load_is_linear(oracle):
    """Under a tiny load the Newton answer approaches the linear solve."""
    m = box_tet10(6, 2, 2, seed=10, jitter=0.1, traction=(0.0, 0.0, -1e-3))
    u, counts = oracle.nh_newton(m, load_steps=1, rtol=1e-10)
    csr = oracle.assemble_csr(m)
    fx = m.fixed_mask()
    ul = oracle.solve_direct(csr, fx, np.zeros(m.n_dofs), oracle.load_vector(m))[0]
    assert np.linalg.norm(u - ul) <= 1e-5 * np.linalg.norm(ul)
    assert counts[0] <= 3
