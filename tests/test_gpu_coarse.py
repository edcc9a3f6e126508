"""Two-level (rigid-body modes on aggregates) preconditioned solves on the GPU.

Bars: FP64 solves within 1e-8 relative L2 of the oracle's direct solve; the
PCG iteration count within 3% (+2) of the oracle's numpy restatement of the
same two-level preconditioner, which pins the aggregation, the mode bases
and the coarse matrix (a wrong A_c still converges, but not in the same
number of iterations); bitwise determinism; mixed precision and learned
blocks with the coarse level; the Newton path with it.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from jfem import mesh as M

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg0(oracle):
    import jfem
    m = M.box_tet10(40, 4, 4, seed=1, jitter=0.1, E_sigma=0.5)
    ctx = jfem.Context(0)
    ctx.upload(m)
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    ud = oracle.solve_direct(csr, m.fixed_mask(), np.zeros(m.n_dofs), f)[0]
    return m, ctx, csr, f, ud


# 243, 88 and 25 aggregates on this mesh
@pytest.mark.parametrize("cn", [32, 64, 256])
def test_two_level_matches_direct_and_oracle_iterations(cfg0, oracle, cn):
    m, ctx, csr, f, ud = cfg0
    say(f"two-level fp64 coarse_nodes={cn}")
    u, _, st = ctx.solve(method=0, op="matfree", rtol=1e-9, coarse_nodes=cn)
    assert st["converged"] == 1
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
    _, it_o = oracle.twolevel_pcg(m, csr, f, cn, rtol=1e-9)
    say(f"  iterations gpu={st['iterations']} oracle={it_o} setup_ms={st['setup_ms']:.2f}")
    assert abs(st["iterations"] - it_o) <= 2 + 0.03 * it_o
    _, _, sb = ctx.solve(method=0, op="matfree", rtol=1e-9)
    assert st["iterations"] * 5 < sb["iterations"]


def test_two_level_csr_and_deterministic(cfg0):
    m, ctx, csr, f, ud = cfg0
    u1, _, s1 = ctx.solve(method=0, op="csr", op_outer="csr", rtol=1e-9, coarse_nodes=128)
    u2, _, s2 = ctx.solve(method=0, op="csr", op_outer="csr", rtol=1e-9, coarse_nodes=128)
    assert s1["iterations"] == s2["iterations"] and np.array_equal(u1, u2)
    assert np.linalg.norm(u1 - ud) / np.linalg.norm(ud) <= 1e-8


def test_two_level_mixed_and_learned(cfg0):
    from jfem.learned import DEFAULT_WEIGHTS
    m, ctx, csr, f, ud = cfg0
    u, _, st = ctx.solve(method=1, op="matfree", op_outer="matfree", rtol=1e-9, inner_rtol=1e-3, coarse_nodes=128)
    say(f"two-level mixed: {st}")
    assert st["converged"] == 1
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
    with open(DEFAULT_WEIGHTS, "rb") as fh:
        ctx.set_learned_weights(fh.read())
    u, _, st = ctx.solve(method=1, op="matfree", op_outer="matfree", precond=1, learned_probe=8, rtol=1e-9,
                         inner_rtol=1e-3, coarse_nodes=128)
    say(f"two-level mixed learned: {st}")
    assert st["converged"] == 1
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8


@pytest.mark.parametrize("gen", [M.box_hex8, M.box_tet4])
def test_two_level_other_elements(oracle, gen):
    import jfem
    m = gen(16, 4, 4, seed=2, jitter=0.1, E_sigma=0.3)
    ctx = jfem.Context(0)
    ctx.upload(m)
    csr = oracle.assemble_csr(m)
    f = oracle.load_vector(m)
    ud = oracle.solve_direct(csr, m.fixed_mask(), np.zeros(m.n_dofs), f)[0]
    u, _, st = ctx.solve(method=0, op="matfree", rtol=1e-10, coarse_nodes=64)
    assert st["converged"] == 1
    assert np.linalg.norm(u - ud) / np.linalg.norm(ud) <= 1e-8
    _, it_o = oracle.twolevel_pcg(m, csr, f, 64, rtol=1e-10)
    assert abs(st["iterations"] - it_o) <= 2 + 0.03 * it_o
    ctx.close()


def test_two_level_newton(oracle):
    import jfem
    m = M.box_tet10(10, 2, 2, seed=3, jitter=0.1, E_sigma=0.3, traction=(0.0, 0.0, -52.5))
    uo, _ = oracle.nh_newton(m, load_steps=5, rtol=1e-10)
    ctx = jfem.Context(0)
    ctx.upload(m)
    for method in (0, 1):
        u, st = ctx.solve_nonlinear(load_steps=5, rtol=1e-10, method=method, coarse_nodes=32)
        say(f"two-level Newton method={method}: {st}")
        assert st["converged"] == 1
        assert np.linalg.norm(u - uo) <= 1e-7 * np.linalg.norm(uo)
    ctx.close()


def test_two_level_config1():
    import jfem
    m = M.box_tet10(160, 16, 16, seed=1, jitter=0.1, E_sigma=0.5)
    ctx = jfem.Context(0)
    ctx.upload(m)
    say("config1 fp64 block Jacobi")
    u0, _, s0 = ctx.solve(method=0, op="matfree", rtol=1e-8)
    for cn, method in ((512, 0), (512, 1)):
        say(f"config1 two-level coarse_nodes={cn} method={method}")
        u, _, st = ctx.solve(method=method, op="matfree", op_outer="matfree", rtol=1e-8, inner_rtol=1e-3,
                             coarse_nodes=cn)
        say(f"  {st}")
        assert st["converged"] == 1
        assert np.linalg.norm(u - u0) / np.linalg.norm(u0) <= 1e-6
        assert st["iterations"] * 5 < s0["iterations"]


_CHILD = r"""
import sys
import numpy as np
sys.path[:0] = [{root!r}, {pkg!r}]
import jfem
from jfem import mesh as M
m = M.box_tet10(40, 4, 4, seed=1, jitter=0.1, E_sigma=0.5)
c = jfem.Context(0)
c.upload(m)
out = {{}}
for meth in (0, 1):
    u, _, st = c.solve(method=meth, op="matfree", op_outer="matfree", rtol=1e-10, inner_rtol=1e-3, coarse_nodes=32)
    out["u%d" % meth] = u
    out["it%d" % meth] = st["iterations"]
np.savez({out!r}, **out)
"""


def test_fused_restriction_matches_separate_restriction(cfg0, tmp_path):
    """Per iteration the coarse right-hand side w = Q' m r is formed inside
    the update kernel, one block per aggregate, from the new residual it
    holds.  Against the restriction by a kernel of its own after a
    node-ordered update (JFEM_FUSED_RESTRICTION=0, a child process): the
    same answer, iteration counts within rounding (3% + 2), in FP64 PCG and
    in the FP32 inner solves."""
    m, ctx, csr, f, ud = cfg0
    res = {}
    for meth in (0, 1):
        say(f"restriction inside the update, method={meth}")
        u, _, st = ctx.solve(method=meth, op="matfree", op_outer="matfree", rtol=1e-10, inner_rtol=1e-3,
                             coarse_nodes=32)
        assert st["converged"] == 1
        res[meth] = (u, st["iterations"])
    out = str(tmp_path / "direct.npz")
    code = _CHILD.format(root=ROOT, pkg=os.path.join(ROOT, "juliafem.jl_amd"), out=out)
    say("separate restriction kernel in a child process (JFEM_FUSED_RESTRICTION=0)")
    env = dict(os.environ, JFEM_FUSED_RESTRICTION="0")
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=120)
    r = np.load(out)
    for meth in (0, 1):
        u, it = res[meth]
        it_d = int(r["it%d" % meth])
        say(f"  method={meth}: iterations fused={it} separate={it_d}")
        assert abs(it - it_d) <= 2 + 0.03 * it_d
        assert np.linalg.norm(u - r["u%d" % meth]) <= 1e-8 * np.linalg.norm(u)
        assert np.linalg.norm(u - ud) <= 1e-8 * np.linalg.norm(ud)


def test_deferred_shared_sums_match_shared_node_kernel(cfg0, tmp_path):
    """In the two-level PCG the update kernel sums the shared nodes' patch
    partials itself (the shared-node kernel of the patch operator does not
    run).  The sums are the same in the same order, so against the solve
    with the shared-node kernel (JFEM_DEFER_SHARED=0, a child process) the
    iterations and the answer are identical, bit for bit, in FP64 PCG and in
    the FP32 inner solves."""
    m, ctx, csr, f, ud = cfg0
    res = {}
    for meth in (0, 1):
        say(f"shared sums in the update, method={meth}")
        u, _, st = ctx.solve(method=meth, op="matfree", op_outer="matfree", rtol=1e-10, inner_rtol=1e-3,
                             coarse_nodes=32)
        assert st["converged"] == 1
        res[meth] = (u, st["iterations"])
    out = str(tmp_path / "combine.npz")
    code = _CHILD.format(root=ROOT, pkg=os.path.join(ROOT, "juliafem.jl_amd"), out=out)
    say("shared-node kernel in a child process (JFEM_DEFER_SHARED=0)")
    env = dict(os.environ, JFEM_DEFER_SHARED="0")
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=120)
    r = np.load(out)
    for meth in (0, 1):
        u, it = res[meth]
        say(f"  method={meth}: iterations deferred={it} kernel={int(r['it%d' % meth])}")
        assert it == int(r["it%d" % meth])
        assert np.array_equal(u, r["u%d" % meth])
        assert np.linalg.norm(u - ud) <= 1e-8 * np.linalg.norm(ud)


_CHILD_AC = r"""
import sys
import numpy as np
sys.path[:0] = [{root!r}, {pkg!r}]
import jfem
from jfem import mesh as M
m = M.{gen}({dims}, seed={seed}, jitter=0.1, E_sigma={sig})
c = jfem.Context(0)
c.upload(m)
A, agg = c.coarse_matrix({cn})
np.savez({out!r}, A=A, agg=agg)
"""


def _block_norms(A, nag):
    B = A.reshape(nag, 6, nag, 6)
    return np.sqrt((B ** 2).sum(axis=(1, 3)))


@pytest.mark.parametrize("gen,dims,seed,sig,cn", [("box_tet10", "40, 4, 4", 1, 0.5, 32),
                                                  ("box_tet10", "40, 4, 4", 1, 0.5, 256),
                                                  ("box_hex8", "16, 4, 4", 2, 0.3, 64)])
def test_coarse_matrix_element_products(oracle, tmp_path, gen, dims, seed, sig, cn):
    """A_c = Q' K Q formed element by element (elements inside one aggregate
    skipped: K_e annihilates the rigid modes) against (1) the same matrix
    by probing K with the modes of distance-2 coloured aggregates
    (JFEM_COARSE_PROBE=1, a child process; the same basis, so entry by
    entry) and (2) the oracle's P' K P with P from a QR of the masked modes
    (another orthonormal basis of each aggregate's span, so the invariants:
    the aggregation bit for bit, the eigenvalues and the Frobenius norm of
    every aggregate-pair block)."""
    import jfem
    m = getattr(M, gen)(*[int(v) for v in dims.split(",")], seed=seed, jitter=0.1, E_sigma=sig)
    ctx = jfem.Context(0)
    ctx.upload(m)
    say(f"coarse matrix {gen} {dims} coarse_nodes={cn}")
    A, agg = ctx.coarse_matrix(cn)
    ctx.close()
    nag = int(agg.max()) + 1
    assert A.shape == (6 * nag, 6 * nag)
    assert np.array_equal(A, A.T)
    s = np.abs(A).max()
    # (1) probing, in a child process
    out = str(tmp_path / "probe.npz")
    code = _CHILD_AC.format(root=ROOT, pkg=os.path.join(ROOT, "juliafem.jl_amd"), gen=gen, dims=dims, seed=seed,
                            sig=sig, cn=cn, out=out)
    subprocess.run([sys.executable, "-c", code], env=dict(os.environ, JFEM_COARSE_PROBE="1"), check=True,
                   timeout=120)
    r = np.load(out)
    assert np.array_equal(r["agg"], agg)
    dp = np.abs(A - r["A"]).max()
    say(f"  n_agg={nag} max|A|={s:.3e} max|A_elem - A_probe|={dp:.2e}")
    assert dp <= 1e-11 * s
    # (2) the oracle
    import scipy.sparse as sp
    assert np.array_equal(agg, oracle.coarse_aggregates(m, cn))
    csr = oracle.assemble_csr(m)
    P, _, _ = oracle.twolevel_operator(m, csr, cn)
    assert P.shape[1] == 6 * nag          # no mode dropped on these meshes
    n = m.n_dofs
    K = sp.csr_matrix((csr[2], csr[1], csr[0]), shape=(n, n))
    D = sp.diags((m.fixed_mask() == 0).astype(np.float64))
    Ao = (P.T @ (D @ K @ D) @ P).toarray()
    ev, evo = np.linalg.eigvalsh(A), np.linalg.eigvalsh(Ao)
    de = np.abs(ev - evo).max() / evo.max()
    db = np.abs(_block_norms(A, nag) - _block_norms(Ao, nag)).max() / s
    say(f"  eigenvalues {evo.min():.3e}..{evo.max():.3e}: rel diff {de:.2e}; block norms rel diff {db:.2e}")
    assert de <= 1e-11 and db <= 1e-11
