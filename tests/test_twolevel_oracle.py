"""Pins the oracle's block-Jacobi and two-level helpers (oracle.block_jacobi,
coarse_aggregates, twolevel_operator, twolevel_pcg) on the CPU.

The checks are properties the mathematics fixes, computed by routes that do
not share the oracle's code: the inverse blocks against the diagonal blocks
read straight from the oracle's CSR; the aggregates against a brute-force
cell partition; the coarse basis against rigid-body fields generated here
(rotations about the origin, not the aggregate centroid); A_c against
Q'KQ formed column by column with the oracle's C SpMV; the two-level PCG
against the banded direct solve.  A swapped index, a dropped rotation mode,
a missing mask or a transposed operand fails one of them.
"""
import numpy as np
import pytest

from jfem.mesh import box_hex8, box_tet10


@pytest.fixture(scope="module")
def small(oracle):
    m = box_tet10(12, 3, 3, seed=5, jitter=0.1, E_sigma=0.5)
    csr = oracle.assemble_csr(m)
    return m, csr


def _diag_block(csr, a):
    rowptr, colind, vals = csr
    B = np.zeros((3, 3))
    for c in range(3):
        s, e = rowptr[3 * a + c], rowptr[3 * a + c + 1]
        for k in range(s, e):
            if 3 * a <= colind[k] < 3 * a + 3:
                B[c, colind[k] - 3 * a] = vals[k]
    return B


def test_block_jacobi_inverts_free_blocks(small, oracle):
    m, csr = small
    fixed = m.fixed_mask()
    Minv = oracle.block_jacobi(csr, fixed).reshape(-1, 3, 3)
    n_checked = 0
    for a in range(m.n_nodes):
        B = _diag_block(csr, a)
        free = fixed[3 * a:3 * a + 3] == 0
        Mi = Minv[a]
        # fixed rows and columns of the inverse are zero
        assert not np.any(Mi[~free, :]) and not np.any(Mi[:, ~free])
        if free.any():
            prod = Mi[np.ix_(free, free)] @ B[np.ix_(free, free)]
            assert np.abs(prod - np.eye(free.sum())).max() <= 1e-12
            n_checked += 1
    # every clamped node (all three DOFs fixed) has a zero block
    assert n_checked == m.n_nodes - m.clamp_nodes.size


def test_block_jacobi_partially_fixed_node(oracle):
    """A node with one fixed component keeps the inverse of its 2x2 free block."""
    m = box_hex8(3, 2, 2, seed=1, jitter=0.1)
    fixed = np.zeros(m.n_dofs, np.uint8)
    fixed[3 * 7 + 1] = 1
    csr = oracle.assemble_csr(m)
    Mi = oracle.block_jacobi(csr, fixed).reshape(-1, 3, 3)[7]
    B = _diag_block(csr, 7)
    keep = [0, 2]
    assert np.allclose(Mi[np.ix_(keep, keep)], np.linalg.inv(B[np.ix_(keep, keep)]), rtol=1e-13, atol=0)
    assert not np.any(Mi[1]) and not np.any(Mi[:, 1])


def test_aggregates_are_the_coordinate_cells(oracle):
    """With every cell above the merge threshold the aggregates are exactly the
    cells of a uniform grid over the bounding box: checked by brute force,
    assigning each node to the half-open interval that holds it."""
    m = box_tet10(8, 2, 2)                     # 17 x 5 x 5 P2 nodes on [0,10]x[0,1]x[0,1]
    agg = oracle.coarse_aggregates(m, 50)
    # about 425 / 50 nodes per cell, cubic cells: 9 slabs along x, one across
    edges = np.linspace(0.0, 10.0, 10)
    slab = np.array([min(8, int(np.searchsorted(edges, x, side="right")) - 1) for x in m.X[:, 0]])
    assert agg.max() + 1 == 9
    # the same partition up to labels
    pairs = {(int(a), int(s)) for a, s in zip(agg, slab)}
    assert len(pairs) == 9


def test_aggregates_merge_small_cells(oracle):
    """Cells below max(8, coarse_nodes/8) nodes join an element-connected
    neighbour: every aggregate ends at least that large, and together they
    still cover every node once."""
    m = box_tet10(10, 3, 3, seed=2, jitter=0.1)
    for cn in (40, 96, 300):
        agg = oracle.coarse_aggregates(m, cn)
        cnt = np.bincount(agg)
        assert agg.shape == (m.n_nodes,) and agg.min() == 0
        assert np.all(cnt > 0)
        assert cnt.min() >= max(8, cn // 8)
        # aggregates are element-connected: each one's nodes form one component
        conn = np.asarray(m.conn)
        for g in range(cnt.size):
            nodes = set(np.flatnonzero(agg == g).tolist())
            start = next(iter(nodes))
            seen, stack = {start}, [start]
            inc = {}
            for e, row in enumerate(conn):
                for a in row:
                    if a in nodes:
                        inc.setdefault(int(a), []).append(e)
            while stack:
                a = stack.pop()
                for e in inc.get(a, []):
                    for b in conn[e]:
                        b = int(b)
                        if b in nodes and b not in seen:
                            seen.add(b)
                            stack.append(b)
            assert seen == nodes


def _rigid_fields(X):
    n = X.shape[0]
    out = []
    for c in range(3):
        v = np.zeros((n, 3))
        v[:, c] = 1.0
        out.append(v.ravel())
    for axis in np.eye(3):
        out.append(np.cross(axis, X).ravel())
    return np.array(out).T                    # n_dofs x 6


@pytest.fixture(scope="module")
def twolevel(small, oracle):
    m, csr = small
    P, Ainv, Minv = oracle.twolevel_operator(m, csr, 128)
    return m, csr, P.toarray(), Ainv, Minv


def test_coarse_basis_orthonormal(twolevel):
    _, _, P, _, _ = twolevel
    assert np.abs(P.T @ P - np.eye(P.shape[1])).max() <= 1e-12


def test_coarse_basis_spans_masked_rigid_modes(twolevel, oracle):
    m, _, P, _, _ = twolevel
    agg = oracle.coarse_aggregates(m, 128)
    fixed = m.fixed_mask() == 1
    R = _rigid_fields(m.X)
    R[fixed] = 0.0
    dof_agg = np.repeat(agg, 3)
    total = 0
    for g in range(agg.max() + 1):
        dofs = np.flatnonzero(dof_agg == g)
        cols = np.flatnonzero(np.abs(P[dofs]).sum(0) > 0)
        # the basis of g lives on g's DOFs only, and on no fixed DOF
        assert np.abs(P[np.setdiff1d(np.arange(m.n_dofs), dofs)][:, cols]).max() == 0
        assert np.abs(P[np.flatnonzero(fixed)][:, cols]).max(initial=0.0) == 0
        Pg, Rg = P[dofs][:, cols], R[dofs]
        # span(P_g) = span(masked rigid fields on g)
        assert np.linalg.matrix_rank(Rg, tol=1e-8 * np.abs(Rg).max()) == cols.size
        resid = Rg - Pg @ (Pg.T @ Rg)
        assert np.abs(resid).max() <= 1e-10 * np.abs(Rg).max()
        total += cols.size
    assert total == P.shape[1]
    # unclamped aggregates carry all six modes
    assert P.shape[1] >= 6 * (agg.max() + 1) - 6 * 2


def test_coarse_matrix_is_galerkin_product(twolevel, oracle):
    """A_c^-1 (the oracle's) times Q'KQ formed here column by column with the
    C SpMV and the Dirichlet mask applied by hand is the identity."""
    m, csr, P, Ainv, _ = twolevel
    fixed = m.fixed_mask() == 1
    KP = np.empty_like(P)
    for j in range(P.shape[1]):
        x = P[:, j].copy()
        x[fixed] = 0.0
        y = oracle.spmv(csr, x)
        y[fixed] = 0.0
        KP[:, j] = y
    Ac = P.T @ KP
    assert np.abs(Ac - Ac.T).max() <= 1e-10 * np.abs(Ac).max()
    assert np.abs(Ainv @ Ac - np.eye(Ac.shape[0])).max() <= 1e-8
    # the coarse correction is the K-orthogonal projector onto range(Q):
    # Q A_c^-1 Q' K v = v for v in range(Q)
    v = P @ np.random.default_rng(1).standard_normal(P.shape[1])
    Kv = oracle.spmv(csr, np.where(fixed, 0.0, v))
    Kv[fixed] = 0.0
    assert np.abs(P @ (Ainv @ (P.T @ Kv)) - v).max() <= 1e-8 * np.abs(v).max()


def test_twolevel_pcg_matches_direct_solve(twolevel, oracle):
    m, csr, _, _, _ = twolevel
    f = oracle.load_vector(m)
    ud, _, _ = oracle.solve_direct(csr, m.fixed_mask(), np.zeros(m.n_dofs), f)
    x, it2 = oracle.twolevel_pcg(m, csr, f, 128, rtol=1e-11)
    assert np.linalg.norm(x - ud) / np.linalg.norm(ud) <= 1e-8
    _, it1, _, rc = oracle.pcg_bj(csr, m.fixed_mask(), np.where(m.fixed_mask() == 1, 0.0, f), rtol=1e-11)
    assert rc == 0
    # the coarse level removes the slow bending modes: far fewer iterations
    assert it2 < 0.6 * it1
