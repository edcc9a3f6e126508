"""jfm_apply_noalloc (the K.u that allocates nothing) against jfm_apply,
through the C ABI, on device vectors from the library's own HIP runtime.

Two calls in a row on different meshes, the second smaller, in FP64 and
FP32, for the matrix-free and the stored-element-matrix operators; then the
error paths (oversized mesh, missing element matrices, unsupported operator).
Bars as for jfm_apply: FP64 within 1e-12, FP32 within 2e-6 of the FP64 result
(max-norm, relative to the largest entry).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from jfem import _lib
from jfem.mesh import box_tet10

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import jfem
    lib = jfem.load()                     # brings in the HIP runtime the library uses
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return jfem, lib, hip


class DeviceVec:
    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), 8 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, 8 * self.n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def noalloc(env, ctx, op, prec, x):
    jfem, lib, hip = env
    xd, yd = DeviceVec(hip, x.size), DeviceVec(hip, x.size)
    try:
        xd.put(x)
        yd.put(np.full(x.size, np.nan))
        assert hip.hipDeviceSynchronize() == 0
        rc = lib.jfm_apply_noalloc(ctx._h, _lib.OPS[op], prec, xd.p, yd.p, None)
        assert hip.hipDeviceSynchronize() == 0
        if rc != 0:
            return rc, None
        return 0, yd.get()
    finally:
        xd.free()
        yd.free()


def meshes():
    return [box_tet10(12, 3, 3, seed=5, jitter=0.1, E_sigma=0.5), box_tet10(6, 2, 2, seed=3, jitter=0.1, E_sigma=0.3)]


@pytest.mark.parametrize("op", ["matfree", "ebe"])
def test_noalloc_matches_apply_two_calls(env, op):
    jfem = env[0]
    ctxs = []
    for k, m in enumerate(meshes()):               # the second mesh is smaller
        ctx = jfem.Context(0)
        ctx.upload(m)
        ctxs.append((m, ctx))
    for m, ctx in ctxs:
        x = np.random.default_rng(m.n_dofs).standard_normal(m.n_dofs)
        ref = ctx.apply(x, op=op, precision=0)    # also builds the element matrices for ebe
        s = np.abs(ref).max()
        for prec, tol in ((0, 1e-12), (1, 2e-6)):
            say(f"noalloc op={op} prec={prec} n_dofs={m.n_dofs}")
            rc, y = noalloc(env, ctx, op, prec, x)
            assert rc == 0, ctx._lib.jfm_last_error(ctx._h)
            assert np.abs(y - ref).max() <= tol * s
            # and against the existing entry point in the same precision
            y_same = ctx.apply(x, op=op, precision=prec)
            assert np.abs(y - y_same).max() <= tol * s
    for _, ctx in ctxs:
        ctx.close()


def test_noalloc_rejects_what_it_cannot_hold(env):
    jfem = env[0]
    m = box_tet10(6, 2, 2)
    ctx = jfem.Context(0)
    ctx.upload(m)
    x = np.ones(m.n_dofs)
    rc, _ = noalloc(env, ctx, "ebe", 0, x)     # element matrices not built yet
    assert rc == -5 and b"element matrices" in ctx._lib.jfm_last_error(ctx._h)
    rc, _ = noalloc(env, ctx, "csr", 0, x)
    assert rc == -7
    ctx.close()
    # 307,200 Tet10 elements need 73.7 MB of FP64 element results > 64 MiB
    big = box_tet10(200, 16, 16)
    ctx = jfem.Context(0)
    say("noalloc oversize mesh")
    ctx.upload_mesh(big.X, big.conn, big.etype, big.E, big.nu)
    rc, _ = noalloc(env, ctx, "matfree", 0, np.zeros(big.n_dofs))
    assert rc == -1 and b"bytes of scratch" in ctx._lib.jfm_last_error(ctx._h)
    # the FP32 form needs half as much and fits
    rc, y = noalloc(env, ctx, "matfree", 1, np.zeros(big.n_dofs))
    assert rc == 0 and not np.any(y)
    ctx.close()
