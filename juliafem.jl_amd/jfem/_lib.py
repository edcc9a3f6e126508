"""ctypes binding of libjfem_hip.so (declarations: include/jfem.h).

There is no CPU fallback: if the HIP library is not built, or no GPU is
visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libjfem_hip.so")

OP_CSR, OP_EBE, OP_MATFREE, OP_TANGENT = 0, 1, 2, 3
FP64, FP32 = 0, 1
SOLVE_PCG, SOLVE_MIXED_IR = 0, 1
PREC_BLOCK_JACOBI, PREC_LEARNED, PREC_USER = 0, 1, 2
NPE = {"Tet10": 10, "Tet4": 4, "Hex8": 8}
NFN = {"Tri6": 6, "Tri3": 3, "Quad4": 4}
OPS = {"csr": OP_CSR, "ebe": OP_EBE, "matfree": OP_MATFREE, "tangent": OP_TANGENT}
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOTSPD", -4: "ENOCONV", -5: "ESTATE", -6: "ENOMEM", -7: "EUNSUPPORTED"}

# Exported symbols (must match include/jfem.h)
SYMBOLS = ("jfm_create", "jfm_destroy", "jfm_last_error", "jfm_get_info", "jfm_upload_mesh",
           "jfm_upload_bcs", "jfm_csr_pattern", "jfm_assemble", "jfm_assemble_tl", "jfm_apply", "jfm_apply_noalloc", "jfm_time_apply", "jfm_coarse_matrix",
           "jfm_solve", "jfm_solve_history", "jfm_precond_set_blocks", "jfm_learned_blob_bytes", "jfm_learned_n_params",
           "jfm_learned_set_weights", "jfm_learned_get_weights", "jfm_learned_features", "jfm_learned_outputs",
           "jfm_learned_grad", "jfm_nl_residual", "jfm_solve_nonlinear", "jfm_postprocess", "jfm_heat_solve")


class JfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("op", C.c_int32), ("op_outer", C.c_int32), ("precond", C.c_int32),
                ("rtol", C.c_double), ("inner_rtol", C.c_double), ("maxit", C.c_int32),
                ("max_outer", C.c_int32), ("check_every", C.c_int32), ("learned_probe", C.c_int32),
                ("inner_maxit", C.c_int32), ("first_rtol", C.c_double), ("coarse_nodes", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer", C.c_int32), ("converged", C.c_int32),
                ("fallback", C.c_int32), ("relres", C.c_double), ("solve_ms", C.c_double),
                ("setup_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NewtonOpts(C.Structure):
    _fields_ = [("load_steps", C.c_int32), ("max_newton", C.c_int32), ("rtol", C.c_double), ("method", C.c_int32),
                ("precond", C.c_int32), ("eta_max", C.c_double), ("maxit_linear", C.c_int32),
                ("learned_probe", C.c_int32), ("inner_rtol", C.c_double), ("inner_maxit", C.c_int32),
                ("check_every", C.c_int32), ("coarse_nodes", C.c_int32)]


class NewtonStats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("linear_iterations", C.c_int32), ("converged", C.c_int32),
                ("fallbacks", C.c_int32), ("relres", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Info(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_el", C.c_int64), ("n_dofs", C.c_int64), ("nnz", C.c_int64),
                ("nnzb", C.c_int64), ("n_faces", C.c_int64), ("npe", C.c_int32), ("affine", C.c_int32),
                ("device_bytes", C.c_int64), ("n_patches", C.c_int64), ("patch_nodes", C.c_int64),
                ("shared_nodes", C.c_int64), ("shared_slots", C.c_int64), ("patch_conn_bytes", C.c_int64),
                ("patch_fp32_factor_rows", C.c_int64)]


_lib = None
_P = C.c_void_p


def load():
    """Load the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          f"make -C juliafem.jl_amd/csrc")
    lib = C.CDLL(LIB_PATH)
    lib.jfm_create.argtypes = [C.c_int, C.POINTER(_P)]
    lib.jfm_destroy.argtypes = [_P]
    lib.jfm_destroy.restype = None
    lib.jfm_last_error.argtypes = [_P]
    lib.jfm_last_error.restype = C.c_char_p
    lib.jfm_get_info.argtypes = [_P, C.POINTER(Info)]
    lib.jfm_upload_mesh.argtypes = [_P, C.c_int64, _P, C.c_int64, C.c_int32, _P, _P, _P, C.c_int32]
    lib.jfm_upload_bcs.argtypes = [_P, _P, _P, C.c_int64, C.c_int32, _P, _P, C.c_int32]
    lib.jfm_csr_pattern.argtypes = [_P, C.POINTER(C.c_int64), _P, _P, C.c_int32]
    lib.jfm_assemble.argtypes = [_P, _P, _P, _P]
    lib.jfm_assemble_tl.argtypes = [_P, C.c_int32, _P, _P, _P, _P]
    lib.jfm_apply.argtypes = [_P, C.c_int32, C.c_int32, _P, _P]
    lib.jfm_apply_noalloc.argtypes = [_P, C.c_int32, C.c_int32, _P, _P, _P]
    lib.jfm_time_apply.argtypes = [_P, C.c_int32, C.c_int32, C.c_int32, _P]
    lib.jfm_coarse_matrix.argtypes = [_P, C.c_int32, C.c_int32, _P, _P, _P]
    lib.jfm_solve.argtypes = [_P, C.POINTER(SolveOpts), _P, _P, _P, C.POINTER(SolveStats)]
    lib.jfm_solve_history.argtypes = [_P, _P, _P, C.c_int32]
    lib.jfm_precond_set_blocks.argtypes = [_P, _P]
    lib.jfm_learned_blob_bytes.argtypes = []
    lib.jfm_learned_blob_bytes.restype = C.c_size_t
    lib.jfm_learned_n_params.argtypes = []
    lib.jfm_learned_n_params.restype = C.c_int32
    lib.jfm_learned_features.argtypes = [_P, _P]
    lib.jfm_learned_outputs.argtypes = [_P, _P, _P]
    lib.jfm_learned_grad.argtypes = [_P, C.c_int32, _P, _P, C.c_int32, C.POINTER(C.c_double), _P]
    lib.jfm_learned_set_weights.argtypes = [_P, _P, C.c_size_t]
    lib.jfm_nl_residual.argtypes = [_P, _P, C.c_double, _P, C.POINTER(C.c_double)]
    lib.jfm_postprocess.argtypes = [_P, _P, _P, _P, _P, _P]
    lib.jfm_heat_solve.argtypes = [_P, _P, _P, C.c_int64, C.c_int32, _P, _P, _P, _P, _P, _P, C.c_int32, C.c_double,
                                   C.c_int32, _P, C.POINTER(SolveStats)]
    lib.jfm_solve_nonlinear.argtypes = [_P, C.POINTER(NewtonOpts), _P, C.POINTER(NewtonStats)]
    lib.jfm_learned_get_weights.argtypes = [_P, _P, C.c_size_t]
    for name in SYMBOLS:
        if name not in ("jfm_destroy", "jfm_last_error", "jfm_learned_blob_bytes", "jfm_learned_n_params"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


class Context:
    """One device context (include/jfem.h jfm_ctx)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _P()
        rc = lib.jfm_create(device, C.byref(h))
        if rc != 0:
            raise JfemError(rc, f"jfm_create(device={device}) failed (no HIP device?)")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.jfm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise JfemError(rc, self._lib.jfm_last_error(self._h).decode())

    # -- upload ------------------------------------------------------------
    def upload_mesh(self, X, conn, etype, E, nu, index_base: int = 0):
        X = np.ascontiguousarray(X, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        n_el = conn.shape[0]
        E = np.ascontiguousarray(np.broadcast_to(E, (n_el,)), dtype=np.float64)
        nu = np.ascontiguousarray(np.broadcast_to(nu, (n_el,)), dtype=np.float64)
        self._check(self._lib.jfm_upload_mesh(self._h, X.shape[0], _p(X), n_el, NPE[etype], _p(conn),
                                               _p(E), _p(nu), index_base))

    def upload_bcs(self, fixed, g=None, faces=None, ftype="Tri6", traction=None, index_base: int = 0):
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
        g = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        if faces is None or len(faces) == 0:
            nf, fc, tr = 0, None, None
        else:
            fc = np.ascontiguousarray(faces, dtype=np.int64)
            nf = fc.shape[0]
            tr = np.ascontiguousarray(np.broadcast_to(np.asarray(traction, dtype=np.float64), (nf, 3)))
        self._check(self._lib.jfm_upload_bcs(self._h, _p(fixed), _p(g), nf, NFN[ftype], _p(fc), _p(tr),
                                              index_base))

    def upload(self, mesh):
        """Mesh + clamp + traction of a jfem.mesh.Mesh."""
        self.upload_mesh(mesh.X, mesh.conn, mesh.etype, mesh.E, mesh.nu)
        self.upload_bcs(mesh.fixed_mask(), None, mesh.faces, mesh.ftype, mesh.face_traction())

    # -- queries -----------------------------------------------------------
    def info(self) -> dict:
        inf = Info()
        self._check(self._lib.jfm_get_info(self._h, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    def csr_pattern(self, base: int = 0):
        nnz = C.c_int64(0)
        self._check(self._lib.jfm_csr_pattern(self._h, C.byref(nnz), None, None, base))
        n = self.info()["n_dofs"]
        rowptr = np.empty(n + 1, np.int64)
        colind = np.empty(nnz.value, np.int64)
        self._check(self._lib.jfm_csr_pattern(self._h, C.byref(nnz), _p(rowptr), _p(colind), base))
        return rowptr, colind

    # -- compute -----------------------------------------------------------
    def assemble(self, u=None, want_vals: bool = True, want_f: bool = True):
        inf = self.info()
        vals = np.empty(inf["nnz"]) if want_vals else None
        f = np.empty(inf["n_dofs"]) if want_f else None
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        self._check(self._lib.jfm_assemble(self._h, _p(uu), _p(vals), _p(f)))
        return vals, f

    def assemble_tl(self, u=None, geometric=True, want_vals=True, want_f=True):
        """Finite-strain assembly: (Km values, Kg values or None, f = f_ext - f_int(u));
        the CSR operator then holds Km + Kg."""
        inf = self.info()
        vals = np.empty(inf["nnz"]) if want_vals else None
        kg = np.empty(inf["nnz"]) if (want_vals and geometric) else None
        f = np.empty(inf["n_dofs"]) if want_f else None
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        self._check(self._lib.jfm_assemble_tl(self._h, int(geometric), _p(uu), _p(vals), _p(kg), _p(f)))
        return vals, kg, f

    def apply(self, x, op="matfree", precision=FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._check(self._lib.jfm_apply(self._h, OPS[op], precision, _p(x), _p(y)))
        return y

    def time_apply(self, op="matfree", precision=FP64, reps: int = 20):
        """(total, first kernel, nodal gather) average ms per K.u."""
        ms = (C.c_double * 3)()
        self._check(self._lib.jfm_time_apply(self._h, OPS[op], precision, reps, ms))
        return tuple(ms)

    def coarse_matrix(self, coarse_nodes: int, op="matfree"):
        """(A_c = Q' K Q before inversion, node aggregates) of the two-level
        preconditioner for coarse_nodes (op "tangent": the neo-Hookean
        tangent at the stored state)."""
        nc = C.c_int64(0)
        self._check(self._lib.jfm_coarse_matrix(self._h, OPS[op], coarse_nodes, C.byref(nc), None, None))
        A = np.empty((nc.value, nc.value))
        agg = np.empty(self.info()["n_nodes"], np.int32)
        self._check(self._lib.jfm_coarse_matrix(self._h, OPS[op], coarse_nodes, C.byref(nc), _p(A), _p(agg)))
        return A, agg

    def solve(self, f=None, method=SOLVE_PCG, op="matfree", op_outer="matfree", precond=PREC_BLOCK_JACOBI,
              rtol=1e-8, inner_rtol=1e-4, maxit=100000, max_outer=20, check_every=16, learned_probe=0,
              inner_maxit=0, first_rtol=0.0, coarse_nodes=0, want_la=False, raise_on_noconv=True):
        n = self.info()["n_dofs"]
        opts = SolveOpts(method, OPS[op], OPS[op_outer], precond, rtol, inner_rtol, maxit, max_outer,
                         check_every, learned_probe, inner_maxit, first_rtol, coarse_nodes)
        st = SolveStats()
        u = np.empty(n)
        la = np.empty(n) if want_la else None
        ff = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        rc = self._lib.jfm_solve(self._h, C.byref(opts), _p(ff), _p(u), _p(la), C.byref(st))
        if rc != 0 and not (rc == -4 and not raise_on_noconv):
            self._check(rc)
        return u, la, st.as_dict()

    def nl_residual(self, u, load_factor=1.0):
        """(R = lf f_ext - f_int(u) on free DOFs, |R|) for the neo-Hookean model;
        u also becomes the state of op="tangent"."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        R = np.empty_like(u)
        rn = C.c_double(0.0)
        self._check(self._lib.jfm_nl_residual(self._h, _p(u), load_factor, _p(R), C.byref(rn)))
        return R, rn.value

    def solve_nonlinear(self, load_steps=5, max_newton=20, rtol=1e-8, method=SOLVE_PCG, precond=PREC_BLOCK_JACOBI,
                        eta_max=0.1, maxit_linear=20000, learned_probe=0, inner_rtol=1e-2, inner_maxit=0,
                        check_every=32, coarse_nodes=0, raise_on_noconv=True):
        """Load-stepped Newton for the compressible neo-Hookean model; (u, stats)."""
        n = self.info()["n_dofs"]
        opts = NewtonOpts(load_steps, max_newton, rtol, method, precond, eta_max, maxit_linear, learned_probe,
                          inner_rtol, inner_maxit, check_every, coarse_nodes)
        st = NewtonStats()
        u = np.empty(n)
        rc = self._lib.jfm_solve_nonlinear(self._h, C.byref(opts), _p(u), C.byref(st))
        if rc != 0 and not (rc == -4 and not raise_on_noconv):
            self._check(rc)
        return u, st.as_dict()

    def postprocess(self, u, gauss=True, nodal=True):
        """Gauss-point strain/stress (n_el, n_gauss, 6) and their least-squares
        nodal fits (n_nodes, 6), ABAQUS order 11, 22, 33, 12, 23, 13."""
        inf = self.info()
        ngp = {10: 4, 4: 1, 8: 8}[inf["npe"]]
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = {}
        if gauss:
            out["strain"] = np.empty((inf["n_el"], ngp, 6))
            out["stress"] = np.empty((inf["n_el"], ngp, 6))
        if nodal:
            out["strain_nodal"] = np.empty((inf["n_nodes"], 6))
            out["stress_nodal"] = np.empty((inf["n_nodes"], 6))
        self._check(self._lib.jfm_postprocess(self._h, _p(u), _p(out.get("strain")), _p(out.get("stress")),
                                              _p(out.get("strain_nodal")), _p(out.get("stress_nodal"))))
        return out

    def heat_solve(self, k, q=None, faces=None, flux=None, h=None, T_ext=None, fixed=None, T_fixed=None,
                   rtol=1e-10, maxit=20000, raise_on_noconv=True):
        """Steady conduction on the uploaded mesh; returns (T per node, stats).
        k, q per element; faces (n_faces, nfn) 0-based with flux / h / T_ext per
        face; fixed (per node, bool) with T_fixed (per node)."""
        inf = self.info()
        N, n_el = inf["n_nodes"], inf["n_el"]
        f64 = lambda a, m: None if a is None else np.ascontiguousarray(np.broadcast_to(a, (m,)), dtype=np.float64)
        kk, qq = f64(k, n_el), f64(q, n_el)
        if faces is None:
            faces = np.zeros((0, 3), np.int64)
        faces = np.ascontiguousarray(faces, dtype=np.int64)
        nf = faces.shape[0]
        fl, hh, te = f64(flux, nf), f64(h, nf), f64(T_ext, nf)
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        tf = f64(T_fixed, N)
        T = np.empty(N)
        st = SolveStats()
        rc = self._lib.jfm_heat_solve(self._h, _p(kk), _p(qq), nf, faces.shape[1], _p(faces) if nf else None,
                                      _p(fl), _p(hh), _p(te), _p(fx), _p(tf), 0, rtol, maxit, _p(T), C.byref(st))
        if rc != 0 and not (rc == -4 and not raise_on_noconv):
            self._check(rc)
        return T, st.as_dict()

    def history(self, max_entries: int = 64):
        """[(cumulative iterations, FP64 relative residual)] of the last solve
        (one entry per Newton iteration after solve_nonlinear)."""
        it = np.zeros(max_entries, np.int32)
        rr = np.zeros(max_entries)
        n = self._lib.jfm_solve_history(self._h, _p(it), _p(rr), max_entries)
        if n < 0:
            self._check(n)
        return list(zip(it[:n].tolist(), rr[:n].tolist()))

    def set_preconditioner_blocks(self, M):
        """User 3x3 blocks (n_nodes x 9) for precond=PREC_USER."""
        M = np.ascontiguousarray(M, dtype=np.float64)
        self._check(self._lib.jfm_precond_set_blocks(self._h, _p(M)))

    def set_learned_weights(self, blob: bytes):
        buf = np.frombuffer(blob, dtype=np.uint8).copy()
        self._check(self._lib.jfm_learned_set_weights(self._h, _p(buf), buf.size))

    def get_learned_weights(self) -> bytes:
        n = self._lib.jfm_learned_blob_bytes()
        buf = np.empty(n, np.uint8)
        self._check(self._lib.jfm_learned_get_weights(self._h, _p(buf), n))
        return buf.tobytes()

    def learned_features(self):
        N = self.info()["n_nodes"]
        f = np.empty((N, 24), np.float32)
        self._check(self._lib.jfm_learned_features(self._h, _p(f)))
        return f

    def learned_outputs(self):
        """(network outputs N x 6, preconditioner blocks N x 9) for the current weights."""
        N = self.info()["n_nodes"]
        o = np.empty((N, 6), np.float32)
        M = np.empty((N, 9))
        self._check(self._lib.jfm_learned_outputs(self._h, _p(o), _p(M)))
        return o, M

    def learned_grad(self, R, E, coarse_nodes=0):
        """(mean loss, gradient) over residual/error pairs (rows of R and E);
        coarse_nodes > 0 trains the blocks as the fine level of the two-level
        preconditioner."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        g = np.empty(self._lib.jfm_learned_n_params(), np.float32)
        loss = C.c_double(0)
        self._check(self._lib.jfm_learned_grad(self._h, R.shape[0], _p(R), _p(E), coarse_nodes, C.byref(loss), _p(g)))
        return loss.value, g
