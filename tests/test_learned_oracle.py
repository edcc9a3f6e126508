"""Pins the oracle's learned-preconditioner restatement: its analytic gradient
against central finite differences (FP64, no BF16 rounding), SPD blocks, and
the block-Jacobi limit of an untrained network.  CPU only."""
import numpy as np
import pytest

from jfem.mesh import box_tet10


@pytest.fixture(scope="module")
def problem(oracle):
    m = box_tet10(6, 2, 2, seed=3, jitter=0.1, E_sigma=0.5)
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    fx = m.fixed_mask()
    Minv = oracle.block_jacobi(csr, fx)
    feat = oracle.learned_features(m, fx)
    us, _, _ = oracle.solve_direct(csr, fx, np.zeros(m.n_dofs), f)
    R, E = [], []
    for k in (3, 10, 30):
        xk, _, _, _ = oracle.pcg_bj(csr, fx, f, rtol=1e-30, maxit=k)
        e = us - xk
        e[fx == 1] = 0
        r = oracle.spmv(csr, e)
        r[fx == 1] = 0
        R.append(r)
        E.append(e)
    return m, csr, fx, Minv, feat, np.array(R), np.array(E)


def random_net(feat, seed, w3=0.3):
    rng = np.random.default_rng(seed)
    return {"W1": rng.standard_normal((64, 24)) * 0.3, "b1": rng.standard_normal(64) * 0.1,
            "W2": rng.standard_normal((64, 64)) * 0.15, "b2": rng.standard_normal(64) * 0.1,
            "W3": rng.standard_normal((6, 64)) * w3 * 0.1, "b3": rng.standard_normal(6) * 0.1,
            "shift": feat.mean(0), "scale": 1.0 / np.maximum(feat.std(0), 1e-6), "eps": 1e-3}


def test_gradient_matches_finite_differences(oracle, problem):
    m, csr, fx, Minv, feat, R, E = problem
    net = random_net(feat, 1)
    loss, grad = oracle.learned_loss_grad(csr, fx, Minv, feat, net, R, E, emulate_bf16=False)
    assert 0.0 <= loss < 1.0
    offsets = {"W1": 0, "b1": 1536, "W2": 1600, "b2": 5696, "W3": 5760, "b3": 6144}
    rng = np.random.default_rng(0)
    checked = 0
    for key, off in offsets.items():
        flat = np.asarray(net[key], dtype=np.float64).ravel()
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            h = 3e-5 * max(1.0, abs(flat[idx]))
            vals = []
            for sgn in (1, -1):
                pert = dict(net)
                arr = flat.copy()
                arr[idx] += sgn * h
                pert[key] = arr.reshape(np.shape(net[key]))
                vals.append(oracle.learned_loss_grad(csr, fx, Minv, feat, pert, R, E, emulate_bf16=False)[0])
            fd = (vals[0] - vals[1]) / (2 * h)
            g = grad[off + idx]
            assert abs(fd - g) <= 2e-6 * max(abs(fd), 1e-2 * np.abs(grad).max()), (key, idx, fd, g)
            checked += 1
    assert checked == 36


def test_blocks_spd_and_block_jacobi_limit(oracle, problem):
    m, csr, fx, Minv, feat, R, E = problem
    _, M = oracle.learned_blocks(feat, random_net(feat, 2, w3=3.0), Minv, fx, emulate_bf16=False)
    free = np.array([not fx[3 * a:3 * a + 3].any() for a in range(m.n_nodes)])
    for a in np.flatnonzero(free):
        B = M[a].reshape(3, 3)
        assert np.allclose(B, B.T) and np.linalg.eigvalsh(B).min() > 0
    # zero last layer: M = (1 + eps) block Jacobi on free nodes, = block Jacobi elsewhere
    net0 = random_net(feat, 2)
    net0["W3"] = np.zeros((6, 64))
    net0["b3"] = np.zeros(6)
    _, M0 = oracle.learned_blocks(feat, net0, Minv, fx, emulate_bf16=False)
    assert np.allclose(M0[free], 1.001 * Minv[free], rtol=1e-12, atol=0)
    assert np.array_equal(M0[~free], Minv[~free])


def test_features_closed_forms(oracle, problem):
    """Normalised diagonal block has trace 3; shares and fractions in range;
    a uniform-E mesh has zero material contrast."""
    m, csr, fx, Minv, feat, R, E = problem
    assert np.allclose(feat[:, 0] + feat[:, 1] + feat[:, 2], 3.0, atol=1e-12)
    assert (feat[:, 7] > 0).all() and (feat[:, 7] <= 1).all()
    assert set(np.unique(feat[:, 13])) <= {0.0, 1.0}
    m2 = box_tet10(4, 2, 2)
    csr2 = oracle.assemble_csr(m2)
    f2 = oracle.learned_features(m2, m2.fixed_mask())
    assert np.allclose(f2[:, 10], 0) and np.allclose(f2[:, 11], 0, atol=1e-12)
    # diagonal block of the features equals the CSR diagonal block
    a = 40
    rp, ci, v = csr2
    D = np.zeros((3, 3))
    for c in range(3):
        s, e = rp[3 * a + c], rp[3 * a + c + 1]
        msk = (ci[s:e] >= 3 * a) & (ci[s:e] < 3 * a + 3)
        D[c, ci[s:e][msk] - 3 * a] = v[s:e][msk]
    Dn = 3 * D / np.trace(D)
    assert np.allclose(f2[a, [0, 1, 2, 3, 4, 5]], [Dn[0, 0], Dn[1, 1], Dn[2, 2], Dn[0, 1], Dn[1, 2], Dn[0, 2]], atol=1e-12)
