"""Seeded structured meshes of the box [0,Lx]x[0,Ly]x[0,Lz] and an ABAQUS reader.

The reference exports `create_structured_box_mesh` but the file that defines
it is absent from the snapshot (src/JuliaFEM.jl:432-434), so the generator is
written here from the mesh description alone:

* Tet10: every hex of an nx*ny*nz grid is split into the 6 Kuhn tetrahedra
  (all sharing the hex's main diagonal), which is conforming across hexes.
  P2 nodes sit on the refined (2nx+1)*(2ny+1)*(2nz+1) grid; every refined
  grid point is a node (vertex, edge, face-diagonal or body-diagonal mid
  node).  Node ids are x-major: id = (ix*(2ny+1) + iy)*(2nz+1) + iz, which
  keeps K banded (half-bandwidth about two x-planes).
* Local node order is the reference Tet10 order
  (src/topology/tetrahedra.jl:73-105, src/basis/lagrange_generated.jl:261-266):
  corners 1-4, then mid nodes of edges (1-2),(2-3),(3-1),(1-4),(2-4),(3-4).
  Every tet is positively oriented (detJ > 0; the reference uses the signed
  detJ, src/basis/math.jl:202).
* Jitter moves interior vertex nodes only, by U(-a*h, a*h) per coordinate;
  mid nodes are placed at edge midpoints afterwards, so elements stay affine
  and boundary faces stay planar.
* Young's modulus is log-normal per element: E = E_median * exp(sigma * z).
* Randomness is counter based (SplitMix64 finaliser over seed, stream and
  index), so every draw is addressable and the mesh does not depend on the
  order in which it is built.

Face sets: ``clamp`` = nodes on x = 0; ``traction`` = surface elements on
x = Lx (Tri6 for Tet10, Tri3 for Tet4, Quad4 for Hex8), outward normal +x.
"""
from __future__ import annotations

import dataclasses
import numpy as np

__all__ = ["Mesh", "box_tet10", "box_tet4", "box_hex8", "read_abaqus_inp",
           "splitmix_uniform", "TET10_EDGES"]

# Mid-node edges of the reference Tet10 (0-based corner indices).
TET10_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))
# Mid-node edges of the reference Tri6.
TRI6_EDGES = ((0, 1), (1, 2), (2, 0))

# The 6 Kuhn tetrahedra of the unit cube: axis permutation -> corner path
# (0,0,0) -> e_p0 -> e_p0+e_p1 -> (1,1,1).
_KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))

_U64 = np.uint64
_GOLDEN = _U64(0x9E3779B97F4A7C15)
_STREAM_MUL = _U64(0xD1B54A32D192ED03)


def _mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 output finaliser (Steele, Lea & Flood 2014), wrapping uint64."""
    z = z.astype(_U64, copy=True)
    with np.errstate(over="ignore"):
        z ^= z >> _U64(30)
        z *= _U64(0xBF58476D1CE4E5B9)
        z ^= z >> _U64(27)
        z *= _U64(0x94D049BB133111EB)
        z ^= z >> _U64(31)
    return z


def splitmix_uniform(seed: int, stream: int, index: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0,1) addressed by (seed, stream, index)."""
    index = np.asarray(index, dtype=_U64)
    with np.errstate(over="ignore"):
        key = _mix64(np.array([_U64(seed & 0xFFFFFFFFFFFFFFFF) + _U64(stream) * _STREAM_MUL], dtype=_U64))[0]
        z = _mix64(key + (index + _U64(1)) * _GOLDEN)
    return (z >> _U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _lognormal(seed: int, n: int, median: float, sigma: float) -> np.ndarray:
    if sigma == 0.0:
        return np.full(n, float(median))
    i = np.arange(n, dtype=np.uint64)
    u1 = 1.0 - splitmix_uniform(seed, 1, 2 * i)          # (0, 1]
    u2 = splitmix_uniform(seed, 1, 2 * i + 1)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return median * np.exp(sigma * z)


@dataclasses.dataclass
class Mesh:
    """Volume mesh plus the boundary sets of the cantilever problem.

    X: (N, 3) float64 coordinates (row n = node n; identical in memory to the
       reference's 3xN column-major geometry layout).
    conn: (n_el, npe) int64, 0-based, reference local node order.
    etype: "Tet10" | "Tet4" | "Hex8".
    E, nu: (n_el,) float64 per-element material.
    faces: (n_faces, nfn) int64 surface elements carrying the traction.
    ftype: "Tri6" | "Tri3" | "Quad4".
    clamp_nodes: (n_clamp,) int64 nodes with all three DOFs fixed.
    """
    X: np.ndarray
    conn: np.ndarray
    etype: str
    E: np.ndarray
    nu: np.ndarray
    faces: np.ndarray
    ftype: str
    clamp_nodes: np.ndarray
    traction: tuple = (0.0, 0.0, -1.0)
    seed: int | None = None
    grid: tuple | None = None

    @property
    def n_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def n_el(self) -> int:
        return int(self.conn.shape[0])

    @property
    def n_dofs(self) -> int:
        return 3 * self.n_nodes

    def fixed_mask(self) -> np.ndarray:
        """uint8 mask over DOFs (3*node + c): 1 where the DOF is clamped."""
        m = np.zeros(self.n_dofs, dtype=np.uint8)
        for c in range(3):
            m[3 * self.clamp_nodes + c] = 1
        return m

    def face_traction(self) -> np.ndarray:
        """(n_faces, 3) traction vector per surface element."""
        return np.tile(np.asarray(self.traction, dtype=np.float64), (self.faces.shape[0], 1))


def _vertex_grid(nx, ny, nz, L, seed, jitter):
    """Vertex coordinates (nx+1, ny+1, nz+1, 3) with interior jitter."""
    h = np.array([L[0] / nx, L[1] / ny, L[2] / nz])
    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    V = np.stack([gi * h[0], gj * h[1], gk * h[2]], axis=-1).astype(np.float64)
    # exact end coordinates (avoid i*h rounding at the far faces)
    V[nx, :, :, 0] = L[0]
    V[:, ny, :, 1] = L[1]
    V[:, :, nz, 2] = L[2]
    if jitter and seed is not None:
        interior = ((gi > 0) & (gi < nx) & (gj > 0) & (gj < ny) & (gk > 0) & (gk < nz))
        # draws are addressed by the refined-grid node id of the vertex
        NY, NZ = 2 * ny + 1, 2 * nz + 1
        nid = ((2 * gi) * NY + 2 * gj) * NZ + 2 * gk
        for c in range(3):
            u = splitmix_uniform(seed, 0, 3 * nid.astype(np.uint64) + np.uint64(c))
            V[..., c] += np.where(interior, (2.0 * u - 1.0) * jitter * h[c], 0.0)
    return V


def box_tet10(nx: int, ny: int, nz: int, L=(10.0, 1.0, 1.0), seed: int | None = None,
              jitter: float = 0.0, E: float = 210e3, nu: float = 0.3,
              E_sigma: float = 0.0, traction=(0.0, 0.0, -1.0)) -> Mesh:
    """Kuhn 6-tet P2 mesh of the box (see module docstring)."""
    NX, NY, NZ = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    V = _vertex_grid(nx, ny, nz, L, seed, jitter)
    ri, rj, rk = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    lo = (ri // 2, rj // 2, rk // 2)
    hi = ((ri + 1) // 2, (rj + 1) // 2, (rk + 1) // 2)
    X = 0.5 * (V[lo] + V[hi])
    X = X.reshape(-1, 3)

    def nid(a, b, c):
        return (a * NY + b) * NZ + c

    hi_, hj_, hk_ = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi_, hj_, hk_ = hi_.ravel(), hj_.ravel(), hk_.ravel()
    nhex = hi_.size
    conn = np.empty((nhex, 6, 10), dtype=np.int64)
    for t, p in enumerate(_KUHN_PERMS):
        e0 = np.zeros(3, dtype=np.int64)
        e0[p[0]] = 1
        e1 = e0.copy()
        e1[p[1]] = 1
        corners = [np.zeros(3, np.int64), e0, e1, np.ones(3, np.int64)]
        sign = np.linalg.det(np.stack([corners[1], corners[2], corners[3]]).astype(float))
        if sign < 0:
            corners[1], corners[2] = corners[2], corners[1]
        # refined offsets of the 10 nodes
        offs = [2 * c for c in corners] + [corners[a] + corners[b] for a, b in TET10_EDGES]
        for a, o in enumerate(offs):
            conn[:, t, a] = nid(2 * hi_ + o[0], 2 * hj_ + o[1], 2 * hk_ + o[2])
    conn = conn.reshape(-1, 10)

    # traction faces on x = Lx: the hex face's Kuhn diagonal runs (1,0,0)-(1,1,1)
    fj, fk = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    fj, fk = fj.ravel(), fk.ravel()
    tris = (((1, 0, 0), (1, 1, 0), (1, 1, 1)), ((1, 0, 0), (1, 1, 1), (1, 0, 1)))
    faces = np.empty((fj.size, 2, 6), dtype=np.int64)
    i0 = nx - 1
    for t, tri in enumerate(tris):
        cr = [np.array(c) for c in tri]
        offs = [2 * c for c in cr] + [cr[a] + cr[b] for a, b in TRI6_EDGES]
        for a, o in enumerate(offs):
            faces[:, t, a] = nid(2 * i0 + o[0], 2 * fj + o[1], 2 * fk + o[2])
    faces = faces.reshape(-1, 6)

    clamp = nid(0, rj[0].ravel(), rk[0].ravel()).astype(np.int64)
    n_el = conn.shape[0]
    Ev = _lognormal(seed if seed is not None else 0, n_el, E, E_sigma)
    return Mesh(X=X, conn=conn, etype="Tet10", E=Ev, nu=np.full(n_el, float(nu)),
                faces=faces, ftype="Tri6", clamp_nodes=np.sort(clamp),
                traction=tuple(traction), seed=seed, grid=(nx, ny, nz))


def box_tet4(nx: int, ny: int, nz: int, L=(10.0, 1.0, 1.0), seed: int | None = None,
             jitter: float = 0.0, E: float = 210e3, nu: float = 0.3,
             E_sigma: float = 0.0, traction=(0.0, 0.0, -1.0)) -> Mesh:
    """Kuhn 6-tet P1 mesh (corner nodes only), x-major node ids."""
    V = _vertex_grid(nx, ny, nz, L, seed, jitter)
    X = V.reshape(-1, 3)
    NY, NZ = ny + 1, nz + 1

    def nid(a, b, c):
        return (a * NY + b) * NZ + c

    hi_, hj_, hk_ = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi_, hj_, hk_ = hi_.ravel(), hj_.ravel(), hk_.ravel()
    conn = np.empty((hi_.size, 6, 4), dtype=np.int64)
    for t, p in enumerate(_KUHN_PERMS):
        e0 = np.zeros(3, dtype=np.int64)
        e0[p[0]] = 1
        e1 = e0.copy()
        e1[p[1]] = 1
        corners = [np.zeros(3, np.int64), e0, e1, np.ones(3, np.int64)]
        if np.linalg.det(np.stack(corners[1:]).astype(float)) < 0:
            corners[1], corners[2] = corners[2], corners[1]
        for a, o in enumerate(corners):
            conn[:, t, a] = nid(hi_ + o[0], hj_ + o[1], hk_ + o[2])
    conn = conn.reshape(-1, 4)
    fj, fk = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    fj, fk = fj.ravel(), fk.ravel()
    tris = (((1, 0, 0), (1, 1, 0), (1, 1, 1)), ((1, 0, 0), (1, 1, 1), (1, 0, 1)))
    faces = np.empty((fj.size, 2, 3), dtype=np.int64)
    for t, tri in enumerate(tris):
        for a, o in enumerate(tri):
            faces[:, t, a] = nid(nx - 1 + o[0], fj + o[1], fk + o[2])
    faces = faces.reshape(-1, 3)
    gj, gk = np.meshgrid(np.arange(NY), np.arange(NZ), indexing="ij")
    clamp = np.sort(nid(0, gj.ravel(), gk.ravel()).astype(np.int64))
    n_el = conn.shape[0]
    Ev = _lognormal(seed if seed is not None else 0, n_el, E, E_sigma)
    return Mesh(X=X, conn=conn, etype="Tet4", E=Ev, nu=np.full(n_el, float(nu)),
                faces=faces, ftype="Tri3", clamp_nodes=clamp, traction=tuple(traction),
                seed=seed, grid=(nx, ny, nz))


def box_hex8(nx: int, ny: int, nz: int, L=(10.0, 1.0, 1.0), seed: int | None = None,
             jitter: float = 0.0, E: float = 210e3, nu: float = 0.3,
             E_sigma: float = 0.0, traction=(0.0, 0.0, -1.0)) -> Mesh:
    """Hex8 grid in the reference Hex8 corner order
    (src/basis/lagrange_generated.jl:289-294), x-major node ids."""
    V = _vertex_grid(nx, ny, nz, L, seed, jitter)
    X = V.reshape(-1, 3)
    NY, NZ = ny + 1, nz + 1

    def nid(a, b, c):
        return (a * NY + b) * NZ + c

    hi_, hj_, hk_ = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi_, hj_, hk_ = hi_.ravel(), hj_.ravel(), hk_.ravel()
    order = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
    conn = np.stack([nid(hi_ + o[0], hj_ + o[1], hk_ + o[2]) for o in order], axis=1).astype(np.int64)
    fj, fk = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    fj, fk = fj.ravel(), fk.ravel()
    quad = ((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1))
    faces = np.stack([nid(nx - 1 + o[0], fj + o[1], fk + o[2]) for o in quad], axis=1).astype(np.int64)
    gj, gk = np.meshgrid(np.arange(NY), np.arange(NZ), indexing="ij")
    clamp = np.sort(nid(0, gj.ravel(), gk.ravel()).astype(np.int64))
    n_el = conn.shape[0]
    Ev = _lognormal(seed if seed is not None else 0, n_el, E, E_sigma)
    return Mesh(X=X, conn=conn, etype="Hex8", E=Ev, nu=np.full(n_el, float(nu)),
                faces=faces, ftype="Quad4", clamp_nodes=clamp, traction=tuple(traction),
                seed=seed, grid=(nx, ny, nz))


_INP_TYPES = {"C3D10": ("Tet10", 10), "C3D4": ("Tet4", 4), "C3D8": ("Hex8", 8)}


def read_abaqus_inp(path: str):
    """Read *NODE and *ELEMENT blocks of an ABAQUS .inp file.

    Mirrors the element-type mapping of src/readers/parse_mesh.jl:26-27
    (C3D10 -> Tet10 with identical local node order).  Returns
    (X (N,3), {elset: (etype, conn (n,npe) 0-based)}); node ids are
    renumbered densely in ascending file id.
    """
    nodes = {}
    elsets = {}
    mode = None
    cur = None
    pending = []
    with open(path) as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line.startswith("**"):
                continue
            if line.startswith("*"):
                key = line.upper()
                mode = None
                if key.startswith("*NODE"):
                    mode = "node"
                elif key.startswith("*ELEMENT"):
                    opts = dict(kv.split("=", 1) for kv in
                                (p.strip() for p in key.split(",")[1:]) if "=" in kv)
                    et = opts.get("TYPE", "")
                    if et in _INP_TYPES:
                        mode = "elem"
                        name = opts.get("ELSET", "ALL")
                        cur = (name, _INP_TYPES[et])
                        elsets.setdefault(name, (_INP_TYPES[et][0], []))
                continue
            if mode == "node":
                parts = [p for p in line.replace(",", " ").split()]
                nodes[int(parts[0])] = [float(v) for v in parts[1:4]]
            elif mode == "elem":
                parts = [p for p in line.replace(",", " ").split() if p]
                pending.extend(int(p) for p in parts)
                npe = cur[1][1]
                while len(pending) >= npe + 1:
                    elsets[cur[0]][1].append(pending[1:npe + 1])
                    pending = pending[npe + 1:]
    ids = np.array(sorted(nodes), dtype=np.int64)
    remap = {nid: i for i, nid in enumerate(ids)}
    X = np.array([nodes[i] for i in ids], dtype=np.float64)
    out = {}
    for name, (et, rows) in elsets.items():
        conn = np.array([[remap[n] for n in r] for r in rows], dtype=np.int64)
        out[name] = (et, conn)
    return X, out


# ------------------------------------------------------------------- gmsh ----
# gmsh element type -> (JuliaFEM type, nodes, permutation gmsh -> reference
# local order).  gmsh's 10-node tetrahedron puts the (3,2) edge node before
# the (3,1) one; the reference order (src/basis/lagrange_generated.jl:261-266)
# has them the other way round, so positions 8 and 9 swap.
_GMSH_TYPES = {
    2: ("Tri3", 3, None), 9: ("Tri6", 6, None), 3: ("Quad4", 4, None),
    4: ("Tet4", 4, None), 11: ("Tet10", 10, (0, 1, 2, 3, 4, 5, 6, 7, 9, 8)), 5: ("Hex8", 8, None),
}


def read_gmsh(path: str):
    """Read an ASCII gmsh 4.1 mesh (the format of src/gmsh_reader.jl).

    Unlike the reference reader, which keeps Tet4 only and takes entity tags
    for physical groups, this one maps entities to their physical groups
    through $Entities and keeps Tri3/Tri6/Quad4/Tet4/Tet10/Hex8.  Returns
    (X (N, 3), {physical name: (etype, conn (n, npe) 0-based)}); node tags are
    renumbered densely in ascending tag order.
    """
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    pos = 0

    def section(name):
        nonlocal pos
        while pos < len(lines) and lines[pos] != name:
            pos += 1
        if pos == len(lines):
            return False
        pos += 1
        return True

    if not section("$MeshFormat"):
        raise ValueError("not a gmsh file")
    ver = lines[pos].split()
    if not ver[0].startswith("4") or ver[1] != "0":
        raise ValueError(f"only ASCII gmsh 4.x is supported (found {lines[pos]})")
    names = {}
    if section("$PhysicalNames"):
        for k in range(int(lines[pos])):
            parts = lines[pos + 1 + k].split(maxsplit=2)
            names[(int(parts[0]), int(parts[1]))] = parts[2].strip('"')
    phys = {}
    pos = 0
    if section("$Entities"):
        counts = [int(v) for v in lines[pos].split()]
        pos += 1
        for dim in range(4):
            for _ in range(counts[dim]):
                parts = lines[pos].split()
                pos += 1
                tag = int(parts[0])
                k = 4 if dim == 0 else 7
                n_ph = int(parts[k])
                phys[(dim, tag)] = [int(v) for v in parts[k + 1:k + 1 + n_ph]]
    pos = 0
    if not section("$Nodes"):
        raise ValueError("no $Nodes section")
    n_blocks, n_nodes = (int(v) for v in lines[pos].split()[:2])
    pos += 1
    tags, coords = [], []
    for _ in range(n_blocks):
        cnt = int(lines[pos].split()[3])
        pos += 1
        tags.extend(int(lines[pos + k]) for k in range(cnt))
        coords.extend([float(v) for v in lines[pos + cnt + k].split()[:3]] for k in range(cnt))
        pos += 2 * cnt
    order = np.argsort(tags)
    tag_sorted = np.asarray(tags, np.int64)[order]
    X = np.asarray(coords, np.float64)[order]
    if not section("$Elements"):
        raise ValueError("no $Elements section")
    n_blocks = int(lines[pos].split()[0])
    pos += 1
    groups = {}
    for _ in range(n_blocks):
        dim, etag, gtype, cnt = (int(v) for v in lines[pos].split())
        pos += 1
        rows = [lines[pos + k].split() for k in range(cnt)]
        pos += cnt
        if gtype not in _GMSH_TYPES:
            continue
        etype, npe, perm = _GMSH_TYPES[gtype]
        conn = np.asarray([[int(v) for v in r[1:1 + npe]] for r in rows], np.int64)
        conn = np.searchsorted(tag_sorted, conn)
        if perm is not None:
            conn = conn[:, list(perm)]
        for ptag in phys.get((dim, etag), []) or [None]:
            # a physical tag without a name falls back to the name filed under
            # the entity tag, the reference reader's reading of such files
            name = names.get((dim, ptag)) or names.get((dim, etag)) or (
                f"entity{dim}_{etag}" if ptag is None else f"physical{dim}_{ptag}")
            if name in groups:
                prev_type, prev = groups[name]
                if prev_type != etype:
                    raise ValueError(f"group {name} mixes {prev_type} and {etype}")
                groups[name] = (etype, np.vstack([prev, conn]))
            else:
                groups[name] = (etype, conn)
    return X, groups


def mesh_from_groups(X, groups, volume: str, clamp: str, load: str | None = None,
                     traction=(0.0, 0.0, -1.0), E: float = 210e3, nu: float = 0.3) -> Mesh:
    """A Mesh from reader output: `volume` group as elements, every node of
    the `clamp` group clamped, the `load` group's faces carrying `traction`."""
    etype, conn = groups[volume]
    if etype not in ("Tet10", "Tet4", "Hex8"):
        raise ValueError(f"{volume} holds {etype}, not volume elements")
    used = np.unique(conn)
    remap = -np.ones(X.shape[0], np.int64)
    remap[used] = np.arange(used.size)
    clamp_nodes = np.unique(groups[clamp][1])
    if load is not None:
        ftype, faces = groups[load]
    else:
        ftype, faces = {"Tet10": "Tri6", "Tet4": "Tri3", "Hex8": "Quad4"}[etype], np.zeros((0, 3), np.int64)
    n = conn.shape[0]
    return Mesh(X=X[used], conn=remap[conn], etype=etype, E=np.full(n, float(E)), nu=np.full(n, float(nu)),
                faces=remap[faces] if faces.size else np.zeros((0, {"Tri6": 6, "Tri3": 3, "Quad4": 4}[ftype]), np.int64),
                ftype=ftype, clamp_nodes=remap[clamp_nodes[np.isin(clamp_nodes, used)]], traction=tuple(traction))


# ------------------------------------------------------------ renumbering ----
def node_graph(mesh: Mesh):
    """Node adjacency (nodes sharing an element, self excluded) in CSR form
    (ptr, idx), each list sorted by node id: the graph G of
    src/graph/graph_ordering.jl."""
    conn = np.asarray(mesh.conn, np.int64)
    npe = conn.shape[1]
    a = np.repeat(conn, npe, axis=1).ravel()
    b = np.tile(conn, (1, npe)).ravel()
    keep = a != b
    key = np.unique(a[keep] * mesh.n_nodes + b[keep])
    src, dst = key // mesh.n_nodes, key % mesh.n_nodes
    ptr = np.searchsorted(src, np.arange(mesh.n_nodes + 1))
    return ptr, dst


def symrcm(ptr, idx, v: int):
    """Reverse Cuthill-McKee from start vertex v, restating symrcm of
    src/graph/graph_ordering.jl:38-90 (0-based, graph in CSR form): breadth
    first from v; each visited vertex appends its unvisited neighbours in
    ascending degree order (a stable sort, as the reference's InsertionSort
    over its adjacency list); the order is reversed.  Returns (perm, invperm,
    degrees, edge, dist) with perm[new] = old.  As in the reference, every
    vertex must be reachable from v."""
    n = ptr.size - 1
    degrees = np.diff(ptr)
    perm = np.empty(n, np.int64)
    perm[0] = v
    visited = np.zeros(n, bool)
    visited[v] = True
    edge = np.zeros(n, np.int64)
    dist = np.zeros(n, np.int64)
    k = 1
    for i in range(n):
        if i >= k:
            raise ValueError("graph is not connected from the start vertex")
        w0 = perm[i]
        adj = idx[ptr[w0]:ptr[w0 + 1]]
        aux = adj[np.argsort(degrees[adj], kind="stable")]
        new = aux[~visited[aux]]
        visited[new] = True
        perm[k:k + new.size] = new
        edge[new] = w0
        dist[new] = dist[w0] + 1
        k += new.size
    perm = perm[::-1].copy()
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    return perm, inv, degrees, edge, dist


def bfs_levels(ptr, idx, v: int):
    """Breadth-first level of every vertex from v (-1: unreachable)."""
    level = -np.ones(ptr.size - 1, np.int64)
    level[v] = 0
    frontier = np.array([v], np.int64)
    d = 0
    while frontier.size:
        d += 1
        starts, ends = ptr[frontier], ptr[frontier + 1]
        nb = idx[np.repeat(starts - np.cumsum(np.r_[0, (ends - starts)[:-1]]), ends - starts)
                 + np.arange((ends - starts).sum())]
        nb = np.unique(nb)
        frontier = nb[level[nb] < 0]
        level[frontier] = d
    return level


def pseudo_peripheral_node(ptr, idx, start: int) -> int:
    """A start vertex of large eccentricity (George and Liu, 1979): from
    `start`, repeatedly move to a minimum-degree vertex of the last BFS level
    while the eccentricity grows."""
    deg = np.diff(ptr)
    v = int(start)
    lv = bfs_levels(ptr, idx, v)
    ecc = lv.max()
    while True:
        last = np.flatnonzero(lv == ecc)
        u = int(last[np.lexsort((last, deg[last]))[0]])
        lu = bfs_levels(ptr, idx, u)
        if lu.max() <= ecc:
            return v
        v, ecc, lv = u, lu.max(), lu


def rcm_order(mesh: Mesh) -> np.ndarray:
    """Reverse Cuthill-McKee node order of the mesh (perm[new] = old): symrcm
    of src/graph/graph_ordering.jl:38-90 from a pseudo-peripheral start vertex.
    The reference takes the start vertex from its caller and needs a
    connected graph; here each connected component (in order of its lowest
    node id) is ordered from its own pseudo-peripheral vertex, found from the
    component's lowest-degree node."""
    ptr, idx = node_graph(mesh)
    n = ptr.size - 1
    deg = np.diff(ptr)
    done = np.zeros(n, bool)
    parts = []
    for s0 in range(n):
        if done[s0]:
            continue
        nodes = np.flatnonzero(bfs_levels(ptr, idx, s0) >= 0)
        done[nodes] = True
        if nodes.size == n:
            sptr, sidx, local = ptr, idx, nodes
        else:                                   # the component's own CSR graph
            remap = -np.ones(n, np.int64)
            remap[nodes] = np.arange(nodes.size)
            cnt = deg[nodes]
            sptr = np.r_[0, np.cumsum(cnt)]
            sidx = remap[np.concatenate([idx[ptr[a]:ptr[a + 1]] for a in nodes])]
            local = nodes
        start = int(np.argmin(np.diff(sptr)))
        v = pseudo_peripheral_node(sptr, sidx, start)
        perm, _, _, _, _ = symrcm(sptr, sidx, v)
        parts.append(local[perm])
    return np.concatenate(parts)


def renumber(mesh: Mesh, perm: np.ndarray) -> Mesh:
    """The same mesh with node `perm[k]` renamed k."""
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return Mesh(X=mesh.X[perm], conn=inv[mesh.conn], etype=mesh.etype, E=mesh.E, nu=mesh.nu,
                faces=inv[mesh.faces] if mesh.faces.size else mesh.faces, ftype=mesh.ftype,
                clamp_nodes=np.sort(inv[mesh.clamp_nodes]), traction=mesh.traction, seed=mesh.seed, grid=None)


def node_bandwidth(mesh: Mesh) -> int:
    """max |a - b| over node pairs sharing an element."""
    c = np.asarray(mesh.conn, np.int64)
    return int((c.max(axis=1) - c.min(axis=1)).max())
