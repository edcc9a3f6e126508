"""ctypes front end of the CPU oracle (jfem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product library
never imports or calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libjfem_oracle.so")
_ETYPE = {"Tet10": 10, "Tet4": 4, "Hex8": 8}
_FTYPE = {"Tri6": 6, "Tri3": 3, "Quad4": 4}

_P = C.c_void_p
_I64 = C.c_int64


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    lib.orc_element_stiffness.restype = C.c_double
    lib.orc_element_stiffness.argtypes = [C.c_int, _P, C.c_double, C.c_double, _P]
    lib.orc_element_internal_force.restype = None
    lib.orc_element_internal_force.argtypes = [C.c_int, _P, _P, C.c_double, C.c_double, _P]
    lib.orc_surface_load.restype = None
    lib.orc_surface_load.argtypes = [C.c_int, _P, _P, _P]
    lib.orc_volume_rule.restype = C.c_int
    lib.orc_volume_rule.argtypes = [C.c_int, _P, _P]
    lib.orc_coo_size.restype = _I64
    lib.orc_coo_size.argtypes = [_I64, C.c_int]
    lib.orc_assemble_coo.restype = C.c_double
    lib.orc_assemble_coo.argtypes = [_P, _I64, C.c_int, _P, _P, _P, _P, _P, _P, C.c_int]
    lib.orc_coo_to_csr.restype = _I64
    lib.orc_coo_to_csr.argtypes = [_I64, _I64, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_load_vector.restype = None
    lib.orc_load_vector.argtypes = [_I64, _P, _I64, C.c_int, _P, _P, _P]
    lib.orc_spmv.restype = None
    lib.orc_spmv.argtypes = [_I64, _P, _P, _P, _P, _P, C.c_int]
    lib.orc_solve_direct.restype = C.c_int
    lib.orc_solve_direct.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_pcg_bj.restype = C.c_int
    lib.orc_pcg_bj.argtypes = [_I64, _P, _P, _P, _P, _P, _P, C.c_double, C.c_int, _P, _P, C.c_int]
    lib.orc_node_to_elem.restype = None
    lib.orc_node_to_elem.argtypes = [_I64, _I64, C.c_int, _P, _P, _P]
    lib.orc_matfree_apply.restype = None
    lib.orc_matfree_apply.argtypes = [_I64, _P, _I64, C.c_int, _P, _P, _P, _P, _P, _P]
    lib.orc_nh_material.restype = C.c_double
    lib.orc_nh_material.argtypes = [_P, C.c_double, C.c_double, _P, _P]
    lib.orc_nh_element.restype = C.c_double
    lib.orc_nh_element.argtypes = [C.c_int, _P, _P, C.c_double, C.c_double, _P, _P]
    lib.orc_nh_assemble.restype = C.c_double
    lib.orc_nh_assemble.argtypes = [_P, _I64, C.c_int, _P, _P, _P, _P, _P, _I64, _P, _P, _P]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def volume_rule(etype: str):
    w = np.zeros(8)
    xi = np.zeros((8, 3))
    n = lib().orc_volume_rule(_ETYPE[etype], _p(w), _p(xi))
    return w[:n].copy(), xi[:n].copy()


def element_stiffness(etype: str, Xe, E: float, nu: float):
    Xe = _f64(Xe)
    nd = 3 * _ETYPE[etype]
    Ke = np.zeros((nd, nd))
    det = lib().orc_element_stiffness(_ETYPE[etype], _p(Xe), E, nu, _p(Ke))
    return Ke, det


def element_internal_force(etype: str, Xe, ue, E: float, nu: float):
    Xe, ue = _f64(Xe), _f64(ue)
    fe = np.zeros(3 * _ETYPE[etype])
    lib().orc_element_internal_force(_ETYPE[etype], _p(Xe), _p(ue), E, nu, _p(fe))
    return fe


def surface_load(ftype: str, Xf, t):
    Xf, t = _f64(Xf), _f64(t)
    fe = np.zeros(3 * _FTYPE[ftype])
    lib().orc_surface_load(_FTYPE[ftype], _p(Xf), _p(t), _p(fe))
    return fe


def assemble_csr(mesh, nthreads: int = 0):
    """K in CSR with sparse(COO) semantics: (rowptr, colind, vals), int64/f64."""
    et = _ETYPE[mesh.etype]
    X, conn = _f64(mesh.X), _i64(mesh.conn)
    E, nu = _f64(mesh.E), _f64(mesh.nu)
    m = int(lib().orc_coo_size(mesh.n_el, et))
    I = np.empty(m, np.int64)
    J = np.empty(m, np.int64)
    V = np.empty(m, np.float64)
    min_det = lib().orc_assemble_coo(_p(X), mesh.n_el, et, _p(conn), _p(E), _p(nu),
                                     _p(I), _p(J), _p(V), nthreads)
    if not min_det > 0:
        raise ValueError(f"inverted element (min detJ = {min_det})")
    n = mesh.n_dofs
    rowptr = np.empty(n + 1, np.int64)
    colind = np.empty(m, np.int64)
    vals = np.empty(m, np.float64)
    perm = np.empty(m, np.int64)
    nnz = int(lib().orc_coo_to_csr(n, m, _p(I), _p(J), _p(V), _p(rowptr), _p(colind), _p(vals), _p(perm)))
    del I, J, V, perm
    return rowptr, colind[:nnz].copy(), vals[:nnz].copy()


def load_vector(mesh):
    f = np.zeros(mesh.n_dofs)
    faces = _i64(mesh.faces)
    tr = _f64(mesh.face_traction())
    lib().orc_load_vector(mesh.n_dofs, _p(_f64(mesh.X)), faces.shape[0], _FTYPE[mesh.ftype],
                          _p(faces), _p(tr), _p(f))
    return f


def spmv(csr, x, nthreads: int = 0):
    rowptr, colind, vals = csr
    x = _f64(x)
    y = np.empty(rowptr.size - 1)
    lib().orc_spmv(rowptr.size - 1, _p(rowptr), _p(colind), _p(vals), _p(x), _p(y), nthreads)
    return y


def solve_direct(csr, fixed, g, f):
    rowptr, colind, vals = csr
    n = rowptr.size - 1
    fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
    g, f = _f64(g), _f64(f)
    u = np.zeros(n)
    la = np.zeros(n)
    bw = C.c_int64(0)
    rc = lib().orc_solve_direct(n, _p(rowptr), _p(colind), _p(vals), _p(fixed), _p(g), _p(f),
                                _p(u), _p(la), C.byref(bw))
    if rc != 0:
        raise RuntimeError(f"oracle direct solve failed (rc={rc})")
    return u, la, int(bw.value)


def solve_direct_lapack(csr, fixed, g, f):
    """The same Dirichlet elimination and direct solve as solve_direct
    (src/solvers.jl:192-216: u[B] = g[B], K[I,I] u[I] = f[I] - K[I,B] g[B],
    la = f - K u on B), factorised by LAPACK's blocked banded Cholesky
    (dpbtrf/dpbtrs through scipy, multithreaded OpenBLAS) so that meshes of
    ~1e5 DOFs with node-plane bandwidth solve in seconds, followed by one
    step of iterative refinement.  Returns (u, la, bandwidth)."""
    from scipy.linalg import cho_solve_banded, cholesky_banded
    rowptr, colind, vals = csr
    n = rowptr.size - 1
    fixed = np.asarray(fixed).astype(bool)
    g = np.where(fixed, _f64(g), 0.0)
    f = _f64(f)
    fmap = -np.ones(n, np.int64)
    free_idx = np.flatnonzero(~fixed)
    fmap[free_idx] = np.arange(free_idx.size)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    keep = (fmap[rows] >= 0) & (fmap[colind] >= 0) & (fmap[colind] <= fmap[rows])
    r, c, v = fmap[rows[keep]], fmap[colind[keep]], vals[keep]
    bw = int((r - c).max()) if r.size else 0
    ab = np.zeros((bw + 1, free_idx.size))
    np.add.at(ab, (r - c, c), v)               # lower form: ab[i - j, j] = K[i, j]
    # right-hand side f[I] - K[I,B] g[B]
    Kg = np.zeros(n)
    kb = fixed[colind]
    np.add.at(Kg, rows[kb], vals[kb] * g[colind[kb]])
    rhs = (f - Kg)[free_idx]
    L = cholesky_banded(ab, lower=True, check_finite=False)
    del ab
    u = g.copy()
    u[free_idx] = cho_solve_banded((L, True), rhs, check_finite=False)

    def K_times(x):
        y = np.zeros(n)
        np.add.at(y, rows, vals * x[colind])
        return y

    # one step of iterative refinement with the same factor (FP64 residual)
    res = (f - K_times(u))[free_idx]
    u[free_idx] += cho_solve_banded((L, True), res, check_finite=False)
    la = np.where(fixed, f - K_times(u), 0.0)
    return u, la, bw


def pcg_bj(csr, fixed, b, rtol=1e-8, maxit=100000, x0=None, nthreads: int = 0):
    rowptr, colind, vals = csr
    n = rowptr.size - 1
    fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    it = C.c_int(0)
    rel = C.c_double(0)
    rc = lib().orc_pcg_bj(n, _p(rowptr), _p(colind), _p(vals), _p(fixed), _p(_f64(b)), _p(x),
                          rtol, maxit, C.byref(it), C.byref(rel), nthreads)
    return x, int(it.value), float(rel.value), rc


def node_to_elem(mesh):
    ptr = np.empty(mesh.n_nodes + 1, np.int64)
    data = np.empty(mesh.n_el * mesh.conn.shape[1], np.int64)
    lib().orc_node_to_elem(mesh.n_nodes, mesh.n_el, mesh.conn.shape[1], _p(_i64(mesh.conn)), _p(ptr), _p(data))
    return ptr, data


def matfree_apply(mesh, u, fixed=None):
    y = np.empty(mesh.n_dofs)
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
    lib().orc_matfree_apply(mesh.n_nodes, _p(_f64(mesh.X)), mesh.n_el, _ETYPE[mesh.etype],
                            _p(_i64(mesh.conn)), _p(_f64(mesh.E)), _p(_f64(mesh.nu)),
                            None if fx is None else _p(fx), _p(_f64(u)), _p(y))
    return y


def basis(etype: str, xi, surface: bool = False):
    """(N, dN) of a reference element at xi."""
    code = (_FTYPE if surface else _ETYPE)[etype]
    dim = 2 if surface else 3
    xi = _f64(xi)
    N = np.zeros(code)
    dN = np.zeros((code, dim))
    lib().orc_basis.argtypes = [C.c_int, C.c_int, _P, _P, _P]
    lib().orc_basis(code, int(surface), _p(xi), _p(N), _p(dN))
    return N, dN


def surface_rule(ftype: str):
    w = np.zeros(4)
    xi = np.zeros((4, 2))
    lib().orc_surface_rule.argtypes = [C.c_int, _P, _P]
    n = lib().orc_surface_rule(_FTYPE[ftype], _p(w), _p(xi))
    return w[:n].copy(), xi[:n].copy()


# ------------------------------------------------- learned preconditioner --
_U16 = np.uint16


def _net_args(net):
    """Network parameters as FP64 arrays.  `net` is a packed weight blob (the
    BF16 values the library uses) or a dict of FP64 arrays W1 b1 W2 b2 W3 b3
    shift scale eps (for finite-difference checks without rounding)."""
    if isinstance(net, (bytes, bytearray)):
        from jfem.learned import from_bf16, unpack_blob
        w = unpack_blob(net)
        w = {**w, **{k: from_bf16(w[k]).astype(np.float64) for k in ("W1", "W2", "W3")}}
    else:
        w = net
    keep = [_f64(np.asarray(w[k], dtype=np.float64).ravel()) for k in
            ("W1", "b1", "W2", "b2", "W3", "b3", "shift", "scale")]
    return keep, float(w["eps"])


def block_jacobi(csr, fixed):
    """Inverse 3x3 diagonal blocks with fixed rows/columns removed (N x 9)."""
    rowptr, colind, vals = csr
    n = rowptr.size - 1
    N = n // 3
    fixed = np.asarray(fixed, dtype=np.uint8)
    Minv = np.zeros((N, 9))
    for a in range(N):
        B = np.zeros((3, 3))
        for c in range(3):
            row = 3 * a + c
            s, e = rowptr[row], rowptr[row + 1]
            cols = colind[s:e]
            m = (cols >= 3 * a) & (cols < 3 * a + 3)
            B[c, cols[m] - 3 * a] = vals[s:e][m]
        if not np.any(B):
            continue
        for c in range(3):
            if fixed[3 * a + c]:
                B[c, :] = 0
                B[:, c] = 0
                B[c, c] = 1
        Mi = np.linalg.inv(B)
        for c in range(3):
            if fixed[3 * a + c]:
                Mi[c, :] = 0
                Mi[:, c] = 0
        Minv[a] = Mi.ravel()
    return Minv


def learned_features(mesh, fixed, faces=None):
    """Node features; `faces` are the loaded surface elements (default: the mesh's)."""
    lib().orc_learned_features.argtypes = [_I64, _P, _I64, C.c_int, _P, _P, _P, _P, _P, _P]
    lib().orc_learned_features.restype = None
    feat = np.zeros((mesh.n_nodes, 24))
    fx = np.ascontiguousarray(fixed, dtype=np.uint8)
    faces = mesh.faces if faces is None else faces
    onface = np.zeros(mesh.n_nodes, np.uint8)
    onface[np.asarray(faces, dtype=np.int64).ravel()] = 1
    lib().orc_learned_features(mesh.n_nodes, _p(_f64(mesh.X)), mesh.n_el, _ETYPE[mesh.etype], _p(_i64(mesh.conn)),
                               _p(_f64(mesh.E)), _p(_f64(mesh.nu)), _p(fx), _p(onface), _p(feat))
    return feat


def learned_blocks(feat, net, Minv, fixed, emulate_bf16=True):
    keep, eps = _net_args(net)
    N = feat.shape[0]
    out = np.zeros((N, 6))
    M = np.zeros((N, 9))
    f = lib().orc_learned_blocks
    f.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_double, _P, _P, _P, _P, C.c_int]
    f.restype = None
    f(N, _p(_f64(feat)), *[_p(k) for k in keep], eps, _p(_f64(Minv)),
      _p(np.ascontiguousarray(fixed, dtype=np.uint8)), _p(out), _p(M), int(emulate_bf16))
    return out, M


def learned_loss_grad(csr, fixed, Minv, feat, net, R, E, emulate_bf16=True, coarse=None):
    """(mean loss, gradient) over the pairs; `coarse` (pairs x n) is added to
    z = M r when the blocks act inside the two-level preconditioner."""
    keep, eps = _net_args(net)
    rowptr, colind, vals = csr
    N = feat.shape[0]
    R = _f64(R)
    E = _f64(E)
    grad = np.zeros(6150)
    f = lib().orc_learned_loss_grad
    f.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_double, C.c_int, _P, _P, _P, _P,
                  C.c_int]
    f.restype = C.c_double
    Cc = None if coarse is None else _f64(coarse)
    loss = f(N, _p(rowptr), _p(colind), _p(vals), _p(np.ascontiguousarray(fixed, dtype=np.uint8)), _p(_f64(Minv)),
             _p(_f64(feat)), *[_p(k) for k in keep], eps, R.shape[0], _p(R), _p(E), _p(Cc) if Cc is not None else None,
             _p(grad), int(emulate_bf16))
    return loss, grad


# ---------------------------------------------------------------- neo-Hookean --
def nh_material(Eg, mu: float, la: float):
    """(psi, S 3x3, D 6x6 Voigt) of the compressible neo-Hookean at the
    Green-Lagrange strain E."""
    Eg = _f64(np.asarray(Eg, np.float64).reshape(9))
    S = np.empty(9)
    D = np.empty(36)
    psi = lib().orc_nh_material(_p(Eg), mu, la, _p(S), _p(D))
    return psi, S.reshape(3, 3), D.reshape(6, 6)


def nh_element(etype: str, Xe, ue, E: float, nu: float, tangent: bool = True):
    """TL element force (3npe), tangent ((3npe)^2 or None) and min det F."""
    npe = _ETYPE[etype]
    fe = np.empty(3 * npe)
    Ke = np.empty((3 * npe, 3 * npe)) if tangent else None
    d = lib().orc_nh_element(npe, _p(_f64(Xe)), _p(_f64(ue)), E, nu, _p(fe), _p(Ke) if tangent else None)
    return fe, Ke, d


def nh_assemble(mesh, u, tangent: bool = True):
    """(f_int(u), K_T(u) as CSR or None, min det F) over the whole mesh."""
    et = _ETYPE[mesh.etype]
    n = mesh.n_dofs
    fint = np.empty(n)
    args = [None, None, None]
    if tangent:
        m = int(lib().orc_coo_size(mesh.n_el, et))
        I, J, V = np.empty(m, np.int64), np.empty(m, np.int64), np.empty(m)
        args = [_p(I), _p(J), _p(V)]
    d = lib().orc_nh_assemble(_p(_f64(mesh.X)), mesh.n_el, et, _p(_i64(mesh.conn)), _p(_f64(mesh.E)),
                              _p(_f64(mesh.nu)), _p(_f64(u)), _p(fint), n, *args)
    if not tangent:
        return fint, None, d
    rowptr = np.empty(n + 1, np.int64)
    colind = np.empty(m, np.int64)
    vals = np.empty(m)
    perm = np.empty(m, np.int64)
    nnz = int(lib().orc_coo_to_csr(n, m, _p(I), _p(J), _p(V), _p(rowptr), _p(colind), _p(vals), _p(perm)))
    return fint, (rowptr, colind[:nnz].copy(), vals[:nnz].copy()), d


def nh_newton(mesh, load_steps: int = 5, rtol: float = 1e-10, max_newton: int = 30):
    """Load-stepped full Newton with direct tangent solves (scipy SuperLU on
    the free DOFs): the CPU answer for the nonlinear path.  Returns
    (u, per-step Newton iteration counts)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    fixed = mesh.fixed_mask().astype(bool)
    free = ~fixed
    f = load_vector(mesh)
    u = np.zeros(mesh.n_dofs)
    counts = []
    for k in range(1, load_steps + 1):
        lf = k / load_steps
        ref = np.linalg.norm(lf * f[free])
        for it in range(max_newton + 1):
            fint, csr, d = nh_assemble(mesh, u, tangent=True)
            if not d > 0:
                raise ValueError("inverted state in the oracle Newton")
            R = lf * f - fint
            rn = np.linalg.norm(R[free])
            if rn <= rtol * ref:
                break
            if it == max_newton:
                raise RuntimeError(f"oracle Newton did not converge at load step {k}: {rn / ref:.3e}")
            K = sp.csr_matrix((csr[2], csr[1], csr[0]), shape=(mesh.n_dofs,) * 2)
            Kff = K[free][:, free].tocsc()
            du = np.zeros(mesh.n_dofs)
            du[free] = sla.spsolve(Kff, R[free])
            u += du
        counts.append(it)
    return u, counts


# ---------------------------------------------- total Lagrangian (SVK) --
def tl_element(etype: str, Xe, ue, E: float, nu: float, geometric: bool = True):
    """Finite-strain element of Problem{Elasticity} with finite_strain = true,
    restated literally from src/problems_elasticity.jl:203-451: at each Gauss
    point gradu = sum u_k dN_k' (grad!), F = I + gradu, the Green-Lagrange
    strain 1/2 (gradu + gradu' + gradu' gradu) in Voigt order with engineering
    shear (to_voigt!, :189-197), BL(F) (:270-289), the isotropic D (:310-317),
    stress_vec = D strain_vec, Km += w BL' D BL (:369-375), with
    geometric_stiffness Kg += w BNL' S3 BNL (:378-404), f_int += w BL' S
    (:406-409); w = weight * detJ in the reference configuration (:248).
    Returns (Km, Kg, f_int)."""
    Xe = np.asarray(Xe, np.float64)
    ue = np.asarray(ue, np.float64).reshape(-1, 3)
    npe = Xe.shape[0]
    nd = 3 * npe
    w_all, xi_all = volume_rule(etype)
    la = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = la
    D[0, 0] = D[1, 1] = D[2, 2] = 2 * mu + la
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    Km = np.zeros((nd, nd))
    Kg = np.zeros((nd, nd))
    fint = np.zeros(nd)
    for w0, xi in zip(w_all, xi_all):
        _, dNr = basis(etype, xi)                  # npe x 3, d/dxi
        J = dNr.T @ Xe                             # J[a][b] = sum_i dN_i/dxi_a X_i,b
        detJ = np.linalg.det(J)
        dN = np.linalg.solve(J, dNr.T)             # 3 x npe, dN[k, i] = dN_i/dX_k
        w = w0 * detJ
        gradu = ue.T @ dN.T                        # gradu[c, k] = sum_i u_i,c dN_i/dX_k
        F = np.eye(3) + gradu
        strain = 0.5 * (gradu + gradu.T + gradu.T @ gradu)
        sv = np.array([strain[0, 0], strain[1, 1], strain[2, 2],
                       2 * strain[0, 1], 2 * strain[1, 2], 2 * strain[0, 2]])
        BL = np.zeros((6, nd))
        for i in range(npe):
            d1, d2, d3 = dN[0, i], dN[1, i], dN[2, i]
            for c in range(3):
                BL[0, 3 * i + c] = F[c, 0] * d1
                BL[1, 3 * i + c] = F[c, 1] * d2
                BL[2, 3 * i + c] = F[c, 2] * d3
                BL[3, 3 * i + c] = F[c, 0] * d2 + F[c, 1] * d1
                BL[4, 3 * i + c] = F[c, 1] * d3 + F[c, 2] * d2
                BL[5, 3 * i + c] = F[c, 2] * d1 + F[c, 0] * d3
        stress = D @ sv
        Km += w * (BL.T @ D @ BL)
        if geometric:
            BNL = np.zeros((9, nd))
            for i in range(npe):
                for c in range(3):
                    BNL[3 * c:3 * c + 3, 3 * i + c] = dN[:, i]
            S = np.array([[stress[0], stress[3], stress[5]],
                          [stress[3], stress[1], stress[4]],
                          [stress[5], stress[4], stress[2]]])
            S3 = np.kron(np.eye(3), S)
            Kg += w * (BNL.T @ S3 @ BNL)
        fint += w * (BL.T @ stress)
    return Km, Kg, fint


def _coo_csr(n, I, J, V):
    m = I.size
    rowptr = np.empty(n + 1, np.int64)
    colind = np.empty(m, np.int64)
    vals = np.empty(m)
    perm = np.empty(m, np.int64)
    nnz = int(lib().orc_coo_to_csr(n, m, _p(_i64(I)), _p(_i64(J)), _p(_f64(V)), _p(rowptr), _p(colind),
                                   _p(vals), _p(perm)))
    return rowptr, colind[:nnz].copy(), vals[:nnz].copy()


def tl_assemble(mesh, u, geometric: bool = True):
    """(Km CSR, Kg CSR or None, f = -f_int(u)) of the whole mesh: the element
    loop above, COO triplets in ascending element order and sparse(COO)
    (src/sparse/sparse.jl:54, 121-132); f_ext is zero (no volume load)."""
    conn = np.asarray(mesh.conn, np.int64)
    n_el, npe = conn.shape
    n = mesh.n_dofs
    U = np.asarray(u, np.float64).reshape(-1, 3)
    dofs = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(n_el, -1)
    nd = dofs.shape[1]
    I = np.repeat(dofs, nd, axis=1).ravel()
    J = np.tile(dofs, (1, nd)).ravel()
    Vm = np.empty((n_el, nd * nd))
    Vg = np.empty((n_el, nd * nd)) if geometric else None
    f = np.zeros(n)
    for e in range(n_el):
        Km, Kg, fi = tl_element(mesh.etype, mesh.X[conn[e]], U[conn[e]], float(mesh.E[e]), float(mesh.nu[e]),
                                geometric)
        Vm[e] = Km.ravel()
        if geometric:
            Vg[e] = Kg.ravel()
        np.add.at(f, dofs[e], -fi)
    Kmc = _coo_csr(n, I, J, Vm.ravel())
    Kgc = _coo_csr(n, I, J, Vg.ravel()) if geometric else None
    return Kmc, Kgc, f


def tl_newton(mesh, load_steps: int = 1, rtol: float = 1e-10, max_newton: int = 30, geometric: bool = True):
    """run!(::Solver{Nonlinear}) on the assembled TL system (src/solvers.jl:
    575-620): per load step, assemble K = Km + Kg and f = lf f_ext - f_int(u),
    solve the eliminated system directly (SuperLU), u += du, until
    |f| <= rtol |lf f_ext| on the free DOFs.  Returns (u, Newton counts)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    fixed = mesh.fixed_mask().astype(bool)
    free = ~fixed
    fext = load_vector(mesh)
    u = np.zeros(mesh.n_dofs)
    counts = []
    n = mesh.n_dofs
    for k in range(1, load_steps + 1):
        lf = k / load_steps
        ref = np.linalg.norm(lf * fext[free])
        for it in range(max_newton + 1):
            Km, Kg, f = tl_assemble(mesh, u, geometric)
            R = lf * fext + f
            if np.linalg.norm(R[free]) <= rtol * ref:
                break
            if it == max_newton:
                raise RuntimeError("oracle TL Newton did not converge")
            K = sp.csr_matrix((Km[2], Km[1], Km[0]), shape=(n, n))
            if Kg is not None:
                K = K + sp.csr_matrix((Kg[2], Kg[1], Kg[0]), shape=(n, n))
            du = np.zeros(n)
            du[free] = sla.spsolve(K[free][:, free].tocsc(), R[free])
            u += du
        counts.append(it)
    return u, counts


# ------------------------------------------------------------- two-level PCG --
def coarse_aggregates(mesh, coarse_nodes: int):
    """Node aggregates as the library forms them: a coordinate grid with about
    `coarse_nodes` nodes per cell, cells under max(8, coarse_nodes / 8) nodes
    merged into their largest element-connected neighbour."""
    X = np.asarray(mesh.X, np.float64)
    N = X.shape[0]
    lo, hi = X.min(axis=0), X.max(axis=0)
    maxe = (hi - lo).max()
    ext = np.maximum(hi - lo, 1e-12 * max(maxe, 1e-300))
    cell = np.cbrt(ext.prod() / max(1.0, N / max(coarse_nodes, 1)))
    nb = np.maximum(1, np.floor(ext / cell + 0.5).astype(np.int64))
    b = np.floor((X - lo) / ext * nb).astype(np.int64)
    b = np.minimum(np.maximum(b, 0), nb - 1)
    key = (b[:, 0] * nb[1] + b[:, 1]) * nb[2] + b[:, 2]
    _, agg = np.unique(key, return_inverse=True)
    n_agg = int(agg.max()) + 1
    conn = np.asarray(mesh.conn, np.int64)
    # aggregate adjacency through elements
    ea = agg[conn]
    pairs = set()
    for row in ea:
        u = np.unique(row)
        for i in u:
            for j in u:
                if i != j:
                    pairs.add((int(i), int(j)))
    adj = [[] for _ in range(n_agg)]
    for i, j in sorted(pairs):
        adj[i].append(j)
    cnt = np.bincount(agg, minlength=n_agg)
    small = max(8, coarse_nodes // 8)
    into = np.arange(n_agg)
    for g in range(n_agg):
        if cnt[g] >= small:
            continue
        best = -1
        for h in adj[g]:
            if cnt[h] >= small and (best < 0 or cnt[h] > cnt[best]):
                best = h
        if best >= 0:
            into[g] = best
    keep = np.where(into == np.arange(n_agg))[0]
    remap = -np.ones(n_agg, np.int64)
    remap[keep] = np.arange(keep.size)
    return remap[into[agg]]


def twolevel_operator(mesh, csr, coarse_nodes: int):
    """(P, A_c^-1, block Jacobi) of the additive two-level preconditioner
    M^-1 r = B r + P A_c^-1 P' r with P the rigid-body modes of every
    aggregate (masked, orthonormalised) and A_c = P' K P."""
    import scipy.sparse as sp
    agg = coarse_aggregates(mesh, coarse_nodes)
    fixed = mesh.fixed_mask().astype(bool)
    X = np.asarray(mesh.X, np.float64)
    N = X.shape[0]
    n_agg = int(agg.max()) + 1
    cent = np.zeros((n_agg, 3))
    np.add.at(cent, agg, X)
    cent /= np.bincount(agg)[:, None]
    d = X - cent[agg]
    rows, cols, vals = [], [], []
    ncol = 0
    order = np.argsort(agg, kind="stable")
    starts = np.searchsorted(agg[order], np.arange(n_agg + 1))
    for g in range(n_agg):
        nodes = order[starts[g]:starts[g + 1]]
        B = np.zeros((3 * nodes.size, 6))
        for c in range(3):
            B[c::3, c] = 1.0
        for k in range(3):
            r = np.cross(np.eye(3)[k][None, :], d[nodes])
            for c in range(3):
                B[c::3, 3 + k] = r[:, c]
        dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        B[fixed[dofs]] = 0.0
        Q, R = np.linalg.qr(B)
        # Householder QR leaves rounding-level entries in the zero rows
        Q[fixed[dofs]] = 0.0
        dg = np.abs(np.diag(R))
        good = dg > 1e-5 * max(dg.max(), 1e-300)
        for j in np.where(good)[0]:
            rows.append(dofs)
            cols.append(np.full(dofs.size, ncol))
            vals.append(Q[:, j])
            ncol += 1
    n = 3 * N
    P = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, ncol))
    K = sp.csr_matrix((csr[2], csr[1], csr[0]), shape=(n, n))
    D = sp.diags((~fixed).astype(np.float64))
    Ac = (P.T @ (D @ K @ D) @ P).toarray()
    Ainv = np.linalg.inv(0.5 * (Ac + Ac.T))
    Minv = block_jacobi(csr, mesh.fixed_mask()).reshape(N, 3, 3)
    return P, Ainv, Minv


def twolevel_pcg(mesh, csr, b, coarse_nodes: int, rtol: float = 1e-8, maxit: int = 20000):
    """PCG with the two-level preconditioner (numpy); returns (x, iterations)."""
    import scipy.sparse as sp
    P, Ainv, Minv = twolevel_operator(mesh, csr, coarse_nodes)
    fixed = mesh.fixed_mask().astype(bool)
    n = b.size
    K = sp.csr_matrix((csr[2], csr[1], csr[0]), shape=(n, n))
    D = sp.diags((~fixed).astype(np.float64))
    Km = D @ K @ D
    b = np.where(fixed, 0.0, b)

    def prec(r):
        return np.einsum("aij,aj->ai", Minv, r.reshape(-1, 3)).ravel() + P @ (Ainv @ (P.T @ r))

    x = np.zeros(n)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    for it in range(1, maxit + 1):
        q = Km @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= rtol * bn:
            return x, it
        z = prec(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxit


# ------------------------------------------------------------ postprocessing --
def lsq_rule(etype: str):
    """Quadrature of the nodal least-squares fit.  The reference integrates it
    with the element's own rule (src/problems_elasticity.jl:555-560), which
    for Tet10 (4 points) and Tet4 (1 point) makes the fit matrix singular: a
    constant field is not recovered (see tests/test_postprocess_oracle.py).
    Tetrahedra use a collapsed (Duffy) Gauss product rule instead, 3 x 3 x 4
    points, exact to degree 4 (the mass entries of Tet10 are degree 4); Hex8
    keeps its 2 x 2 x 2 rule, which is exact for its mass entries."""
    if etype == "Hex8":
        return volume_rule(etype)
    ga, wa = np.polynomial.legendre.leggauss(3)
    gc, wc = np.polynomial.legendre.leggauss(4)
    ga, wa, gc, wc = 0.5 * (ga + 1), 0.5 * wa, 0.5 * (gc + 1), 0.5 * wc
    pts, wts = [], []
    for a, w1 in zip(ga, wa):
        for b, w2 in zip(ga, wa):
            for c, w3 in zip(gc, wc):
                pts.append([a * (1 - b) * (1 - c), b * (1 - c), c])
                wts.append(w1 * w2 * w3 * (1 - b) * (1 - c) ** 2)
    return np.array(wts), np.array(pts)


def _strain_stress_at(mesh, u, xi):
    conn = np.asarray(mesh.conn, np.int64)
    Xe = np.asarray(mesh.X)[conn]
    ue = np.asarray(u, np.float64).reshape(-1, 3)[conn]
    E, nu = np.asarray(mesh.E), np.asarray(mesh.nu)
    la = (E * nu / ((1 + nu) * (1 - 2 * nu)))[:, None]
    mu = (E / (2 * (1 + nu)))[:, None]
    strain = np.empty((conn.shape[0], len(xi), 6))
    stress = np.empty_like(strain)
    detJ = np.empty((conn.shape[0], len(xi)))
    for q in range(len(xi)):
        _, dN = basis(mesh.etype, xi[q])
        J = np.einsum("ia,eib->eab", dN, Xe)          # J[a][b] = sum_i dN_i/dxi_a X_i,b
        detJ[:, q] = np.linalg.det(J)
        g = np.einsum("eab,ib->eia", np.linalg.inv(J), dN)
        G = np.einsum("eic,eia->eca", ue, g)            # G[c][a] = du_c / dX_a
        ev = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 2, 2], 0.5 * (G[:, 0, 1] + G[:, 1, 0]),
                       0.5 * (G[:, 1, 2] + G[:, 2, 1]), 0.5 * (G[:, 0, 2] + G[:, 2, 0])], axis=1)
        tr = ev[:, :3].sum(axis=1, keepdims=True)
        strain[:, q] = ev
        stress[:, q] = 2 * mu * ev + np.hstack([la * tr] * 3 + [np.zeros_like(tr)] * 3)
    return strain, stress, detJ


def post_gauss(mesh, u):
    """Gauss-point strain and stress (n_el, n_gauss, 6), ABAQUS order 11, 22,
    33, 12, 23, 13 with tensor shear: eps = (grad u + grad u')/2, linear
    isotropic stress (src/problems_elasticity.jl:520-548)."""
    _, xi = volume_rule(mesh.etype)
    strain, stress, _ = _strain_stress_at(mesh, u, xi)
    return strain, stress


def lsq_fit(mesh, u, field: str = "strain", rule=None):
    """Nodal least-squares fit of the strain or stress of u: A x = b with
    A = sum w N N', b = sum w f N, w = weight * detJ, solved by a sparse direct
    method (src/problems_elasticity.jl:550-583), on the points of `rule`
    (default lsq_rule)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    w, xi = lsq_rule(mesh.etype) if rule is None else rule
    strain, stress, detJ = _strain_stress_at(mesh, u, xi)
    fgp = strain if field == "strain" else stress
    conn = np.asarray(mesh.conn, np.int64)
    n_el, npe = conn.shape
    N_nodes = mesh.X.shape[0]
    rows, cols, vals = [], [], []
    b = np.zeros((N_nodes, 6))
    for q in range(len(w)):
        Nq, _ = basis(mesh.etype, xi[q])
        wq = w[q] * detJ[:, q]
        rows.append(np.repeat(conn, npe, axis=1).ravel())
        cols.append(np.tile(conn, (1, npe)).ravel())
        vals.append((wq[:, None, None] * np.outer(Nq, Nq)[None]).ravel())
        np.add.at(b, conn.ravel(), (wq[:, None, None] * Nq[None, :, None] * fgp[:, q, None, :]).reshape(-1, 6))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N_nodes, N_nodes))
    A = 0.5 * (A + A.T)
    return np.column_stack([sla.spsolve(A.tocsc(), b[:, k]) for k in range(6)])


def mass_volume(mesh):
    """sum of the consistent mass entries = the body's volume (a check)."""
    w, xi = lsq_rule(mesh.etype)
    conn = np.asarray(mesh.conn, np.int64)
    Xe = np.asarray(mesh.X)[conn]
    tot = 0.0
    for q in range(w.size):
        Nq, dN = basis(mesh.etype, xi[q])
        tot += (w[q] * np.linalg.det(np.einsum("ia,eib->eab", dN, Xe))).sum() * Nq.sum() ** 2
    return tot


# ---------------------------------------------------------------------- heat --
def heat_system(mesh, k, q=None, faces=None, ftype=None, flux=None, h=None, T_ext=None):
    """(K, f) of steady conduction (src/problems_heat.jl:36-113): volume
    Ke += w detJ k dN'dN, fe += w detJ q N; faces fe += w detJ (g + h T_ext) N,
    Ke += w detJ h N N'.  K as scipy CSR, one DOF per node."""
    import scipy.sparse as sp
    w, xi = volume_rule(mesh.etype)
    conn = np.asarray(mesh.conn, np.int64)
    Xe = np.asarray(mesh.X)[conn]
    n_el, npe = conn.shape
    N = mesh.X.shape[0]
    k = np.broadcast_to(np.asarray(k, np.float64), (n_el,))
    rows, cols, vals = [], [], []
    f = np.zeros(N)
    for qp in range(w.size):
        Nq, dN = basis(mesh.etype, xi[qp])
        J = np.einsum("ia,eib->eab", dN, Xe)
        wd = w[qp] * np.linalg.det(J)
        g = np.einsum("eab,ib->eia", np.linalg.inv(J), dN)
        Ke = np.einsum("e,eia,eja->eij", wd * k, g, g)
        rows.append(np.repeat(conn, npe, axis=1).ravel())
        cols.append(np.tile(conn, (1, npe)).ravel())
        vals.append(Ke.ravel())
        if q is not None:
            qq = np.broadcast_to(np.asarray(q, np.float64), (n_el,))
            np.add.at(f, conn.ravel(), (wd * qq)[:, None].repeat(npe, 1).ravel() * np.tile(Nq, n_el))
    if faces is not None and len(faces):
        faces = np.asarray(faces, np.int64)
        nf, nfn = faces.shape
        Xf = np.asarray(mesh.X)[faces]
        ws, us = surface_rule(ftype)
        gg = np.zeros(nf) if flux is None else np.broadcast_to(np.asarray(flux, np.float64), (nf,))
        hh = np.zeros(nf) if h is None else np.broadcast_to(np.asarray(h, np.float64), (nf,))
        tt = np.zeros(nf) if T_ext is None else np.broadcast_to(np.asarray(T_ext, np.float64), (nf,))
        for qp in range(ws.size):
            Nq, dN = basis(ftype, us[qp], surface=True)
            Jf = np.einsum("ia,eib->eab", dN, Xf)
            area = np.linalg.norm(np.cross(Jf[:, 0], Jf[:, 1]), axis=1)
            wq = ws[qp] * area
            np.add.at(f, faces.ravel(), ((wq * (gg + hh * tt))[:, None] * Nq[None, :]).ravel())
            rows.append(np.repeat(faces, nfn, axis=1).ravel())
            cols.append(np.tile(faces, (1, nfn)).ravel())
            vals.append(((wq * hh)[:, None, None] * np.outer(Nq, Nq)[None]).ravel())
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    return K, f


def heat_solve_direct(K, f, fixed, T_fixed):
    """Dirichlet elimination and a sparse direct solve (as src/solvers.jl:192-216)."""
    import scipy.sparse.linalg as sla
    fixed = np.asarray(fixed, bool)
    T = np.where(fixed, T_fixed, 0.0)
    free = ~fixed
    Kc = K.tocsc()
    b = f[free] - Kc[free][:, fixed] @ T[fixed]
    T[free] = sla.spsolve(Kc[free][:, free], b)
    return T
