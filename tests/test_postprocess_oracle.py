"""Pins the oracle's strain/stress recovery (src/problems_elasticity.jl:520-594).  CPU only.

An affine displacement has constant strain sym(A) at every Gauss point and
its least-squares nodal fit is exact; a quadratic displacement on Tet10 has
a linear strain field, which the fit reproduces at the nodes; the fit's mass
entries sum to the volume and its rule integrates monomials exactly.  The
reference's own rule (the element's 4 points for Tet10) leaves the fit
matrix singular -- shown here -- so the fit uses a richer rule.
"""
import numpy as np
import pytest

from jfem.mesh import box_hex8, box_tet10

from test_oracle import lame


@pytest.mark.parametrize("gen", [box_tet10, box_hex8])
def test_affine_field(oracle, gen):
    m = gen(4, 2, 2, seed=3, jitter=0.15)
    A = np.array([[1.0, 0.3, -0.2], [0.1, -0.5, 0.4], [0.25, 0.05, 0.7]]) * 1e-3
    u = (m.X @ A.T + [0.1, -0.2, 0.3]).ravel()
    strain, stress = oracle.post_gauss(m, u)
    S = 0.5 * (A + A.T)
    ev = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[1, 2], S[0, 2]])
    assert np.abs(strain - ev).max() <= 1e-15
    la, mu = lame(210e3, 0.3)
    sv = 2 * mu * ev + np.r_[[la * np.trace(S)] * 3, [0, 0, 0]]
    assert np.abs(stress - sv).max() <= 1e-10 * np.abs(sv).max()
    assert np.abs(oracle.lsq_fit(m, u, "strain") - ev).max() <= 1e-13
    assert np.abs(oracle.lsq_fit(m, u, "stress") - sv).max() <= 1e-9 * np.abs(sv).max()


def test_quadratic_field_tet10(oracle):
    m = box_tet10(3, 2, 2, seed=2, jitter=0.1)
    x, y, z = m.X.T
    u = np.stack([x * y, z ** 2, x * z], axis=1) * 1e-3
    nodal = oracle.lsq_fit(m, u.ravel(), "strain")
    # u = (xy, z^2, xz): eps_xx = y, eps_yy = 0, eps_zz = x, eps_xy = x/2, eps_yz = z, eps_xz = z/2
    exact = np.stack([y, 0 * x, x, 0.5 * x, z, 0.5 * z], axis=1) * 1e-3
    assert np.abs(nodal - exact).max() <= 1e-12


def test_lsq_rule_exactness(oracle):
    w, xi = oracle.lsq_rule("Tet10")
    assert w.sum() == pytest.approx(1 / 6, rel=1e-14)
    # int over the unit tet of x^a y^b z^c = a! b! c! / (a + b + c + 3)!
    from math import factorial
    for a, b, c in [(4, 0, 0), (0, 4, 0), (0, 0, 4), (2, 2, 0), (1, 1, 2), (1, 2, 1), (3, 0, 1)]:
        exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
        assert (w * xi[:, 0] ** a * xi[:, 1] ** b * xi[:, 2] ** c).sum() == pytest.approx(exact, rel=1e-13)


def test_element_rule_fit_is_singular_for_tet10(oracle):
    """With the element's own 4-point rule (the reference's choice) the fit
    matrix of a Tet10 mesh is singular: a constant strain is not recovered."""
    m = box_tet10(4, 2, 2, seed=3, jitter=0.15)
    A = np.diag([1.0, 2.0, 3.0]) * 1e-3
    u = (m.X @ A.T).ravel()
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        nodal = oracle.lsq_fit(m, u, "strain", rule=oracle.volume_rule("Tet10"))
    assert not np.allclose(nodal[:, 0], 1e-3, rtol=1e-6)


def test_mass_sums_to_volume(oracle):
    m = box_tet10(4, 2, 2, seed=5, jitter=0.1)
    assert oracle.mass_volume(m) == pytest.approx(10.0, rel=1e-12)
