"""Mesh input formats and renumbering (host side).  CPU only.

The gmsh fixtures are the reference's own test data
(test/testdata/cantilever_beam.msh, test/tutorials/01_fundamentals/
reading_gmsh_meshes.msh), read as text.  The Tet10 node-order fix-up is
checked on a one-element file written here in gmsh's order: every
reference mid node must sit on its edge's midpoint.
"""
import os

import numpy as np
import pytest

from jfem.mesh import (TET10_EDGES, Mesh, box_tet10, mesh_from_groups, node_bandwidth, node_graph, rcm_order,
                       read_gmsh, renumber, symrcm)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def tet_volumes(X, conn):
    P = X[conn[:, :4]]
    return np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0]) / 6.0


def test_gmsh_cantilever_fixture():
    X, g = read_gmsh(os.path.join(GOLDEN, "cantilever_beam.msh"))
    assert X.shape == (190, 3)
    assert {k: (v[0], v[1].shape[0]) for k, v in g.items()} == {
        "FixedEnd": ("Tri3", 14), "PressureSurface": ("Tri3", 86), "Bulk": ("Tet4", 434)}
    vol = tet_volumes(X, g["Bulk"][1])
    assert vol.min() > 0 and vol.sum() == pytest.approx(10.0, rel=1e-12)
    assert np.all(X[np.unique(g["FixedEnd"][1]), 0] == 0.0)
    assert np.all(X[np.unique(g["PressureSurface"][1]), 2] == 1.0)
    # the top faces tile the 10 x 1 top surface
    F = X[g["PressureSurface"][1]]
    area = 0.5 * np.linalg.norm(np.cross(F[:, 1] - F[:, 0], F[:, 2] - F[:, 0]), axis=1)
    assert area.sum() == pytest.approx(10.0, rel=1e-12)


def test_gmsh_quad_fixture():
    X, g = read_gmsh(os.path.join(GOLDEN, "reading_gmsh_meshes.msh"))
    etype, conn = g["DOMAIN"]
    assert etype == "Quad4" and conn.shape == (10, 4) and X.shape[0] == 18


def test_gmsh_tet10_order(tmp_path):
    corners = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0], [0, 0, 5]], float)
    gmsh_edges = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]   # gmsh's 10-node tetrahedron
    nodes = np.vstack([corners] + [0.5 * (corners[a] + corners[b]) for a, b in gmsh_edges])
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", "1", '3 7 "Solid"',
             "$EndPhysicalNames", "$Entities", "0 0 0 1", "1 0 0 0 2 3 5 1 7 0", "$EndEntities",
             "$Nodes", f"1 10 1 10", "3 1 0 10"]
    lines += [str(k + 1) for k in range(10)]
    lines += [" ".join(repr(float(v)) for v in p) for p in nodes]
    lines += ["$EndNodes", "$Elements", "1 1 1 1", "3 1 11 1", "1 " + " ".join(str(k + 1) for k in range(10)),
              "$EndElements"]
    path = tmp_path / "tet10.msh"
    path.write_text("\n".join(lines) + "\n")
    X, g = read_gmsh(str(path))
    etype, conn = g["Solid"]
    assert etype == "Tet10"
    P = X[conn[0]]
    for k, (a, b) in enumerate(TET10_EDGES):
        assert np.allclose(P[4 + k], 0.5 * (P[a] + P[b]))


def test_gmsh_mesh_solves(oracle):
    X, g = read_gmsh(os.path.join(GOLDEN, "cantilever_beam.msh"))
    m = mesh_from_groups(X, g, "Bulk", "FixedEnd", "PressureSurface", traction=(0.0, 0.0, -1.0))
    f = oracle.load_vector(m)
    assert f.reshape(-1, 3).sum(axis=0) == pytest.approx([0.0, 0.0, -10.0], abs=1e-12)
    csr = oracle.assemble_csr(m)
    u, la, _ = oracle.solve_direct(csr, m.fixed_mask(), np.zeros(m.n_dofs), f)
    assert u.reshape(-1, 3)[:, 2].min() < 0
    # la = f - K u on the clamped DOFs sums to the applied load (src/solvers.jl:212-213)
    assert la.reshape(-1, 3).sum(axis=0) == pytest.approx([0.0, 0.0, -10.0], abs=1e-8)


def test_rcm_renumbering(oracle):
    m = box_tet10(8, 2, 2, seed=3, jitter=0.1, E_sigma=0.4)
    rng = np.random.default_rng(0)
    shuffled = renumber(m, rng.permutation(m.n_nodes))
    perm = rcm_order(shuffled)
    r = renumber(shuffled, perm)
    assert node_bandwidth(r) < node_bandwidth(shuffled) / 4
    # the physics does not depend on the numbering
    solve = lambda mm: oracle.solve_direct(oracle.assemble_csr(mm), mm.fixed_mask(), np.zeros(mm.n_dofs),
                                           oracle.load_vector(mm))[0].reshape(-1, 3)
    u0, ur = solve(shuffled), solve(r)
    assert np.abs(ur - u0[perm]).max() <= 1e-8 * np.abs(u0).max()


def test_symrcm_known_answer():
    """symrcm (src/graph/graph_ordering.jl:38-90) worked by hand on a small
    graph: from vertex 0 the neighbours are visited in ascending degree, ties
    in the adjacency list's order, and the order is reversed."""
    edges = [(0, 1), (0, 2), (0, 3), (1, 4), (1, 5), (3, 4)]
    n = 6
    adj = [sorted({b for a_, b in edges if a_ == a} | {a_ for a_, b in edges if b == a}) for a in range(n)]
    ptr = np.r_[0, np.cumsum([len(a) for a in adj])]
    idx = np.concatenate([np.array(a, np.int64) for a in adj])
    perm, inv, deg, edge, dist = symrcm(ptr, idx, 0)
    # degrees 3,3,1,2,2,1: from 0 the order is 2 (1), 3 (2), 1 (3); then 3 adds 4, 1 adds 5
    assert perm.tolist() == [5, 4, 1, 3, 2, 0]
    assert np.array_equal(inv[perm], np.arange(n))
    assert deg.tolist() == [3, 3, 1, 2, 2, 1]
    assert dist.tolist() == [0, 1, 1, 1, 2, 2]
    assert [edge[k] for k in (1, 2, 3, 4, 5)] == [0, 0, 0, 3, 1]
    with pytest.raises(ValueError):
        symrcm(np.array([0, 1, 2, 2]), np.array([1, 0]), 0)   # vertex 2 unreachable


def test_rcm_is_a_permutation_and_handles_components():
    a = box_tet10(4, 2, 2, seed=1)
    b = box_tet10(3, 1, 1, seed=2)
    two = Mesh(X=np.vstack([a.X, b.X + 20.0]), conn=np.vstack([a.conn, b.conn + a.n_nodes]), etype="Tet10",
               E=np.r_[a.E, b.E], nu=np.r_[a.nu, b.nu], faces=a.faces, ftype="Tri6", clamp_nodes=a.clamp_nodes)
    shuffled = renumber(two, np.random.default_rng(4).permutation(two.n_nodes))
    perm = rcm_order(shuffled)
    assert np.array_equal(np.sort(perm), np.arange(two.n_nodes))
    r = renumber(shuffled, perm)
    assert node_bandwidth(r) < node_bandwidth(shuffled) / 3
    # each component occupies one contiguous range of the new ids
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    first = np.isin(np.arange(two.n_nodes), np.unique(shuffled.conn[:a.n_el]))
    ids = np.sort(inv[np.flatnonzero(first)])
    assert ids[-1] - ids[0] + 1 == ids.size
