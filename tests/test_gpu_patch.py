"""The patch form of the matrix-free operator (patch.hip), through the C ABI.

The elements are cut into Morton-ordered patches of 480; patch-interior
nodes are summed inside a workgroup and shared nodes by a second kernel.
Checked here beyond the parity suite (which runs every matrix-free product
through it):
  * the structure's invariant: every node is either interior to exactly one
    patch or listed once among the shared nodes (nodes of no element
    included), and there is one patch per 480 elements;
  * K.u against the oracle's restatement (1e-12 FP64, 2e-6 FP32) on meshes
    whose element count is not a multiple of 480 and that hold nodes of no
    element, with and without the row mask (reactions);
  * the two-phase operator (JFEM_PATCH=0, a separate process): the same
    product up to the order of the nodal sums, the same PCG answer;
  * both connectivity widths of the patch kernel (16-bit patch-relative and
    32-bit) and both forms of its FP32 factors (folded, stored) against the
    oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from jfem import mesh as M

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATCH = 480  # elements per patch (kPatch in patch.hip)

MESHES = {
    "tet10_13x3x2": lambda: M.box_tet10(13, 3, 2, seed=7, jitter=0.1, E_sigma=0.4),   # 468 elements, one patch
    "tet10_17x4x3": lambda: M.box_tet10(17, 4, 3, seed=9, jitter=0.1, E_sigma=0.4),   # 1,224 elements
    "tet10_40x4x4": lambda: M.box_tet10(40, 4, 4),                                   # 3,840, unused lattice nodes
    "hex8_9x5x4": lambda: M.box_hex8(9, 5, 4, seed=3, jitter=0.1, E_sigma=0.3),
    "tet4_11x4x3": lambda: M.box_tet4(11, 4, 3, seed=1, jitter=0.1, E_sigma=0.3),
}


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def jf():
    import jfem
    jfem.load()
    return jfem


@pytest.mark.parametrize("name", sorted(MESHES))
def test_structure_and_apply(jf, oracle, name):
    m = MESHES[name]()
    say(f"patch apply {name}: nodes={m.n_nodes} el={m.n_el}")
    ctx = jf.Context(0)
    ctx.upload(m)
    x = np.random.default_rng(11).standard_normal(m.n_dofs)
    y64 = ctx.apply(x, op="matfree")
    inf = ctx.info()
    assert inf["n_patches"] == (m.n_el + PATCH - 1) // PATCH
    interior = inf["patch_nodes"] - inf["shared_slots"]
    assert interior + inf["shared_nodes"] == m.n_nodes
    assert inf["shared_slots"] >= 2 * (inf["shared_nodes"] - (m.n_nodes - len(np.unique(m.conn))))
    fixed = m.fixed_mask()
    ref = oracle.matfree_apply(m, x, fixed)
    scale = np.abs(ref).max()
    assert np.abs(y64 - ref).max() <= 1e-12 * scale
    assert np.all(y64[fixed == 1] == 0.0)
    y32 = ctx.apply(x, op="matfree", precision=1)
    assert np.abs(y32 - ref).max() <= 2e-6 * scale
    # rows of nodes of no element are zero, not stale
    used = np.zeros(m.n_nodes, bool)
    used[np.unique(m.conn)] = True
    assert np.all(y64.reshape(-1, 3)[~used] == 0.0)
    # bitwise repeatable
    assert np.array_equal(ctx.apply(x, op="matfree"), y64)


def test_connectivity_width(jf, oracle):
    """The patch kernel reads 16-bit connectivity relative to each patch's
    lowest node id when every patch's ids span at most 65,536, else the
    32-bit ids.  A 120x8x8 Tet10 mesh (69,649 nodes) takes the 16-bit form
    in its natural numbering and the 32-bit form with its node ids shuffled;
    both products match the oracle's."""
    from jfem.mesh import renumber
    m = M.box_tet10(120, 8, 8, seed=5, jitter=0.1, E_sigma=0.4)
    sh = renumber(m, np.random.default_rng(4).permutation(m.n_nodes))
    for name, mm, width in (("natural", m, 2), ("shuffled", sh, 4)):
        ctx = jf.Context(0)
        ctx.upload(mm)
        x = np.random.default_rng(12).standard_normal(mm.n_dofs)
        y64 = ctx.apply(x, op="matfree")
        y32 = ctx.apply(x, op="matfree", precision=1)
        inf = ctx.info()
        say(f"{name} numbering: {mm.n_nodes} nodes, patch connectivity {inf['patch_conn_bytes']} bytes")
        assert inf["patch_conn_bytes"] == width
        ref = oracle.matfree_apply(mm, x, mm.fixed_mask())
        scale = np.abs(ref).max()
        assert np.abs(y64 - ref).max() <= 1e-12 * scale
        assert np.abs(y32 - ref).max() <= 2e-6 * scale
        ctx.close()


def test_fp32_factor_forms(jf, oracle):
    """With one Poisson's ratio for every element, the FP32 affine Tet10
    patch kernel reads the inverse Jacobians with the shear factor folded in
    (9 rows and lambda / mu as an argument); with per-element ratios, the 11
    stored factors.  Both FP32 products match the oracle's."""
    import dataclasses
    m = M.box_tet10(17, 4, 3, seed=9, jitter=0.1, E_sigma=0.4)
    nu = np.random.default_rng(5).uniform(0.2, 0.35, m.n_el)
    mv = dataclasses.replace(m, nu=nu)
    for name, mm, rows in (("uniform nu", m, 9), ("per-element nu", mv, 11)):
        ctx = jf.Context(0)
        ctx.upload(mm)
        x = np.random.default_rng(13).standard_normal(mm.n_dofs)
        y32 = ctx.apply(x, op="matfree", precision=1)
        y64 = ctx.apply(x, op="matfree")
        inf = ctx.info()
        say(f"{name}: FP32 factor rows {inf['patch_fp32_factor_rows']}")
        assert inf["patch_fp32_factor_rows"] == rows
        ref = oracle.matfree_apply(mm, x, mm.fixed_mask())
        scale = np.abs(ref).max()
        assert np.abs(y64 - ref).max() <= 1e-12 * scale
        assert np.abs(y32 - ref).max() <= 2e-6 * scale
        ctx.close()


def test_reactions_use_unmasked_rows(jf, oracle):
    """la = f - K u on the clamped rows needs the operator without the row
    mask (src/solvers.jl:212-213).  K annihilates rigid translations, so at
    equilibrium the reactions sum to the total applied load."""
    m = M.box_tet10(13, 3, 2, seed=7, jitter=0.1, E_sigma=0.4)
    ctx = jf.Context(0)
    ctx.upload(m)
    u, la, st = ctx.solve(method=0, op="matfree", rtol=1e-10, want_la=True)
    f = oracle.load_vector(m)
    assert np.allclose(la.reshape(-1, 3).sum(0), f.reshape(-1, 3).sum(0), atol=1e-8 * np.abs(f).sum())
    assert np.all(la[m.fixed_mask() == 0] == 0.0)


_CHILD = r"""
import json, sys
import numpy as np
sys.path[:0] = [{root!r}, {pkg!r}]
import jfem
from jfem import mesh as M
m = M.box_tet10(13, 3, 2, seed=7, jitter=0.1, E_sigma=0.4)
c = jfem.Context(0)
c.upload(m)
x = np.random.default_rng(11).standard_normal(m.n_dofs)
y = c.apply(x, op="matfree")
u, _, st = c.solve(method=0, op="matfree", rtol=1e-10)
np.savez({out!r}, y=y, u=u, it=st["iterations"], patches=c.info()["n_patches"])
"""


def test_same_as_two_phase_operator(jf, tmp_path):
    m = M.box_tet10(13, 3, 2, seed=7, jitter=0.1, E_sigma=0.4)
    ctx = jf.Context(0)
    ctx.upload(m)
    x = np.random.default_rng(11).standard_normal(m.n_dofs)
    y = ctx.apply(x, op="matfree")
    u, _, st = ctx.solve(method=0, op="matfree", rtol=1e-10)
    assert ctx.info()["n_patches"] > 0
    out = str(tmp_path / "two_phase.npz")
    code = _CHILD.format(root=ROOT, pkg=os.path.join(ROOT, "juliafem.jl_amd"), out=out)
    say("two-phase operator in a child process (JFEM_PATCH=0)")
    env = dict(os.environ, JFEM_PATCH="0")
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=120)
    r = np.load(out)
    assert int(r["patches"]) == 0
    assert np.abs(y - r["y"]).max() <= 1e-13 * np.abs(y).max()
    # block-Jacobi PCG runs ~950 iterations here, past the loss of
    # orthogonality: rounding-level differences (the order of the nodal sums
    # and of p.q) move the count by a few percent, not the answer
    assert abs(st["iterations"] - int(r["it"])) <= 0.03 * int(r["it"]) + 2
    assert np.linalg.norm(u - r["u"]) <= 1e-9 * np.linalg.norm(u)
