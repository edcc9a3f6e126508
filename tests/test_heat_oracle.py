"""Pins the oracle's heat conduction restatement (src/problems_heat.jl).  CPU only.

A bar [0,10] x [0,1] x [0,1] with insulated sides has closed-form answers
that the elements reproduce at the nodes: a linear profile between fixed
ends, the parabola q x (L - x) / 2k for a uniform source (Tet10 holds it
exactly), a linear profile under an end flux g, and the linear profile of an
end with convection h to T_ext.  Without convection the conduction matrix
annihilates constants.
"""
import numpy as np
import pytest

from jfem.mesh import box_hex8, box_tet4, box_tet10

L, K = 10.0, 3.5


def ends(m):
    x = m.X[:, 0]
    return np.isclose(x, 0.0), np.isclose(x, L)


@pytest.mark.parametrize("gen", [box_tet10, box_hex8, box_tet4])
def test_linear_between_fixed_ends(oracle, gen):
    m = gen(6, 2, 2, seed=1, jitter=0.1)
    left, right = ends(m)
    Kh, f = oracle.heat_system(m, K)
    T = oracle.heat_solve_direct(Kh, f, left | right, np.where(right, 7.0, 2.0))
    assert np.abs(T - (2.0 + 0.5 * m.X[:, 0])).max() <= 1e-11
    assert np.abs(Kh @ np.ones(m.n_nodes)).max() <= 1e-10 * abs(Kh).max()
    assert abs(Kh - Kh.T).max() <= 1e-13 * abs(Kh).max()


def test_uniform_source_parabola_tet10(oracle):
    m = box_tet10(8, 2, 2, seed=2, jitter=0.1)
    left, right = ends(m)
    Kh, f = oracle.heat_system(m, K, q=4.0)
    T = oracle.heat_solve_direct(Kh, f, left | right, np.zeros(m.n_nodes))
    x = m.X[:, 0]
    assert np.abs(T - 4.0 * x * (L - x) / (2 * K)).max() <= 1e-10


@pytest.mark.parametrize("gen,ftype", [(box_tet10, "Tri6"), (box_hex8, "Quad4")])
def test_end_flux_and_convection(oracle, gen, ftype):
    m = gen(6, 2, 2, seed=3, jitter=0.1)
    left, _ = ends(m)
    x = m.X[:, 0]
    # flux g into the x = L face: T = T0 + g x / k
    Kh, f = oracle.heat_system(m, K, faces=m.faces, ftype=ftype, flux=1.5)
    T = oracle.heat_solve_direct(Kh, f, left, np.full(m.n_nodes, 2.0))
    assert np.abs(T - (2.0 + 1.5 * x / K)).max() <= 1e-10
    # convection h to T_ext at x = L: T = T0 + (T_ext - T0) h x / (k + h L)
    h, Te = 0.8, 30.0
    Kh, f = oracle.heat_system(m, K, faces=m.faces, ftype=ftype, h=h, T_ext=Te)
    T = oracle.heat_solve_direct(Kh, f, left, np.full(m.n_nodes, 2.0))
    assert np.abs(T - (2.0 + (Te - 2.0) * h * x / (K + h * L))).max() <= 1e-10
