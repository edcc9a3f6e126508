"""Meshes read from the reference's own files and renumbered meshes, solved
on the GPU through the C ABI and checked against the oracle.

* test/testdata/cantilever_beam.msh (gmsh 4.1: Tet4 bulk, Tri3 clamp and
  pressure faces), read by jfem.mesh.read_gmsh: pattern, K, the Tri3 load
  and the solve;
* a Tet10 cantilever whose node ids are shuffled, then renumbered by
  reverse Cuthill-McKee (jfem.mesh.rcm_order, src/graph/graph_ordering.jl):
  the same displacements, node for node, as the oracle's solve of the
  original numbering.
"""
import os
import sys

import numpy as np
import pytest

from jfem.mesh import box_tet10, mesh_from_groups, node_bandwidth, rcm_order, read_gmsh, renumber

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


def test_gmsh_cantilever_solves(jf, oracle):
    X, g = read_gmsh(os.path.join(GOLDEN, "cantilever_beam.msh"))
    m = mesh_from_groups(X, g, "Bulk", "FixedEnd", "PressureSurface", traction=(0.0, 0.0, -1.0))
    ctx = jf.Context(0)
    say(f"gmsh cantilever: {m.n_nodes} nodes, {m.n_el} Tet4, {m.faces.shape[0]} Tri3 faces")
    ctx.upload(m)
    vals, f = ctx.assemble()
    rowptr, colind = ctx.csr_pattern()
    csr = oracle.assemble_csr(m)
    fr = oracle.load_vector(m)
    assert np.array_equal(rowptr, csr[0]) and np.array_equal(colind, csr[1])
    assert np.abs(vals - csr[2]).max() <= 1e-12 * np.abs(csr[2]).max()
    assert np.abs(f - fr).max() <= 1e-14 * np.abs(fr).max()
    assert f.reshape(-1, 3).sum(0) == pytest.approx([0.0, 0.0, -10.0], abs=1e-12)
    ud, lad, _ = oracle.solve_direct(csr, m.fixed_mask(), np.zeros(m.n_dofs), fr)
    for method in (0, 1):
        u, la, st = ctx.solve(method=method, rtol=1e-10, want_la=True)
        assert st["converged"] == 1
        assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
        fx = m.fixed_mask() == 1
        assert np.abs(la[fx] - lad[fx]).max() <= 1e-6 * np.abs(lad).max()


def test_shuffled_and_rcm_numbering(jf, oracle):
    m = box_tet10(20, 4, 4, seed=6, jitter=0.1, E_sigma=0.5)
    csr = oracle.assemble_csr(m)
    ud, _, _ = oracle.solve_direct_lapack(csr, m.fixed_mask(), np.zeros(m.n_dofs), oracle.load_vector(m))
    Ud = ud.reshape(-1, 3)
    shuffle = np.random.default_rng(3).permutation(m.n_nodes)
    sh = renumber(m, shuffle)
    rcm = rcm_order(sh)
    r = renumber(sh, rcm)
    say(f"bandwidth natural {node_bandwidth(m)}, shuffled {node_bandwidth(sh)}, rcm {node_bandwidth(r)}")
    assert node_bandwidth(r) < node_bandwidth(sh) / 10
    for name, mm, old in (("shuffled", sh, shuffle), ("rcm", r, shuffle[rcm])):
        ctx = jf.Context(0)
        ctx.upload(mm)
        for method, extra in ((0, {}), (1, dict(coarse_nodes=256, inner_rtol=1e-2))):
            u, _, st = ctx.solve(method=method, rtol=1e-9, **extra)
            err = np.linalg.norm(u.reshape(-1, 3) - Ud[old]) / np.linalg.norm(Ud)
            say(f"  {name} method={method}: iterations={st['iterations']} err={err:.2e}")
            assert st["converged"] == 1 and err <= 1e-8
        ctx.close()
