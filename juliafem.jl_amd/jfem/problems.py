"""Host-side mirror of the reference's problem / assembly / solver interface
for the elasticity path, backed by libjfem_hip.so.

Names follow the reference (Julia's `!` dropped): ``assemble`` is
``assemble!(problem, time)`` (src/assembly/assembly.jl:31-70 ->
src/problems_elasticity.jl:65-81), ``solve`` is ``solve!(solver)``
(src/solvers.jl:273-344) and ``run`` is ``run!(analysis)`` for a linear
analysis (src/solvers.jl:641-657).  Error behaviour mirrors the reference:
assembling a non-empty assembly or a problem without elements warns and
returns; unsupported element types raise.

Differences, by design:
* element fields are arrays over an element set, not per-element
  dictionaries, so a million-element problem never builds Python objects
  per element;
* the assembled K arrives as the CSR matrix that ``sparse(COO)`` would
  produce (src/sparse/sparse.jl:54), not as COO triplets; K is structurally
  symmetric, so this is also its CSC pattern;
* ``solve`` replaces the LDL' factorisation by preconditioned CG on the
  GPU (FP64, or FP32 inner iterations inside FP64 refinement) to a relative
  residual, and returns (u, la) like the reference.
"""
from __future__ import annotations

import dataclasses
import warnings

import numpy as np

from . import _lib

VOLUME_TYPES = ("Tet10", "Tet4", "Hex8")
SURFACE_TYPES = ("Tri6", "Tri3", "Quad4")
# element types the reference knows but the 3-D continuum kernel rejects
# (src/problems_elasticity.jl:83, 510-518)
KNOWN_TYPES = VOLUME_TYPES + SURFACE_TYPES + ("Seg2", "Seg3", "Poi1", "Quad8", "Quad9", "Pyr5",
                                              "Wedge6", "Wedge15", "Hex20", "Hex27")


class Elasticity:
    """Problem properties (src/problems_elasticity.jl:36-46)."""

    def __init__(self):
        self.formulation = "continuum"
        self.finite_strain = False
        self.geometric_stiffness = False
        self.store_fields = []


class Dirichlet:
    """Boundary problem properties (src/problems_dirichlet.jl), nodal collocation."""

    def __init__(self):
        self.formulation = "incremental"
        self.variational = False


@dataclasses.dataclass
class ElementSet:
    etype: str
    conn: np.ndarray              # (n, nodes per element), 0-based node ids
    fields: dict


class Assembly:
    """K, f of a field problem; C1, C2, g of a boundary problem
    (src/assembly/problems.jl:14-40)."""

    def __init__(self):
        self.empty()

    def empty(self):
        self.K = None
        self.Kg = None
        self.f = None
        self.C1 = None
        self.C2 = None
        self.g = None
        self.fixed = None

    def isempty(self) -> bool:
        return self.K is None and self.f is None and self.fixed is None


class Problem:
    """Problem{P} (src/assembly/problems.jl:95-179) for P = Elasticity or Dirichlet."""

    def __init__(self, ptype, name: str, dim: int = 3, parent_field_name: str | None = None):
        if dim != 3:
            raise ValueError("the elasticity path is 3-D (dim = 3)")
        self.properties = ptype()
        self.name = name
        self.dim = dim
        self.parent_field_name = parent_field_name
        self.elements: list[ElementSet] = []
        self.X: np.ndarray | None = None          # nodal "geometry"
        self.u: np.ndarray | None = None          # "displacement" field (n_dofs)
        self.assembly = Assembly()
        self._ctx = None

    @property
    def kind(self) -> str:
        return type(self.properties).__name__

    def set_geometry(self, X):
        self.X = np.ascontiguousarray(X, dtype=np.float64)

    def add_elements(self, etype: str, conn, **fields):
        """Add an element set.  Fields (scalar or one value per element):
        Elasticity volume: youngs_modulus, poissons_ratio;
        Elasticity surface: displacement_traction_force (3-vector);
        Dirichlet: displacement_1/2/3 (prescribed component values)."""
        if etype not in KNOWN_TYPES:
            raise ValueError(f"unknown element type {etype}")
        conn = np.asarray(conn, dtype=np.int64)
        if conn.ndim == 1:
            conn = conn[:, None]
        self.elements.append(ElementSet(etype, conn, dict(fields)))

    def context(self):
        if self._ctx is None:
            self._ctx = _lib.Context(0)
        return self._ctx

    # convenience: the cantilever problem of a generated mesh
    @classmethod
    def from_mesh(cls, mesh, name="body"):
        p = cls(Elasticity, name, 3)
        p.set_geometry(mesh.X)
        p.add_elements(mesh.etype, mesh.conn, youngs_modulus=mesh.E, poissons_ratio=mesh.nu)
        p.add_elements(mesh.ftype, mesh.faces, displacement_traction_force=np.asarray(mesh.traction))
        bc = cls(Dirichlet, "fixed", 3, "displacement")
        bc.add_elements("Poi1", mesh.clamp_nodes, displacement_1=0.0, displacement_2=0.0, displacement_3=0.0)
        return p, bc


def _per_element(v, n):
    v = np.asarray(v, dtype=np.float64)
    return np.broadcast_to(v, (n,)) if v.ndim == 0 else v.reshape(n)


def _elasticity_arrays(problem: Problem):
    vol = [s for s in problem.elements if s.etype in VOLUME_TYPES]
    bad = [s.etype for s in problem.elements if s.etype not in VOLUME_TYPES + SURFACE_TYPES]
    if bad:
        # src/problems_elasticity.jl:510-518
        raise RuntimeError(f"Tried to assemble unsupported elements of type {bad[0]} to 3d continuum problem.")
    if len({s.etype for s in vol}) > 1:
        raise RuntimeError("one volume element type per problem on the GPU path")
    if problem.X is None:
        raise RuntimeError("problem has no geometry")
    conn = np.vstack([s.conn for s in vol])
    E = np.concatenate([_per_element(s.fields.get("youngs_modulus"), s.conn.shape[0]) for s in vol])
    nu = np.concatenate([_per_element(s.fields.get("poissons_ratio"), s.conn.shape[0]) for s in vol])
    surf = [s for s in problem.elements if s.etype in SURFACE_TYPES and "displacement_traction_force" in s.fields]
    if len({s.etype for s in surf}) > 1:
        raise RuntimeError("one surface element type per problem on the GPU path")
    if surf:
        faces = np.vstack([s.conn for s in surf])
        tr = np.vstack([np.broadcast_to(np.asarray(s.fields["displacement_traction_force"], float), (s.conn.shape[0], 3))
                        for s in surf])
        ftype = surf[0].etype
    else:
        faces, tr, ftype = np.zeros((0, 6), np.int64), np.zeros((0, 3)), "Tri6"
    return vol[0].etype, conn, E, nu, faces, ftype, tr


def assemble(problem: Problem, time: float = 0.0):
    """assemble!(problem, time): K (CSR, sparse(COO) pattern) and f = f_ext - f_int(u).
    With finite_strain the K is the material tangent Km(u) and, with
    geometric_stiffness, assembly.Kg receives Kg(u) (src/problems_elasticity.jl:
    203-451, 441-447)."""
    if not problem.assembly.isempty():
        warnings.warn("Problem assembly is not empty before assembling. This is probably causing "
                      "unexpected results. To remove old assembly, use `empty!(problem.assembly)`")
        return None
    if not problem.elements:
        warnings.warn("There is no elements defined in problem. Before assembling a problem, elements "
                      "must be added using `add_elements!(problem, elements)`.")
        return None
    if isinstance(problem.properties, Dirichlet):
        return _assemble_dirichlet(problem, time)
    props = problem.properties
    if props.formulation != "continuum":
        raise RuntimeError(f"formulation {props.formulation} is not on the GPU path (3-D continuum only)")
    if props.geometric_stiffness and not props.finite_strain:
        raise RuntimeError("geometric stiffness of the small-strain model is not on the GPU path")
    etype, conn, E, nu, faces, ftype, tr = _elasticity_arrays(problem)
    ctx = problem.context()
    ctx.upload_mesh(problem.X, conn, etype, E, nu)
    n = 3 * problem.X.shape[0]
    ctx.upload_bcs(np.zeros(n, np.uint8), None, faces, ftype, tr)
    kg = None
    if props.finite_strain:
        vals, kg, f = ctx.assemble_tl(u=problem.u, geometric=props.geometric_stiffness)
    else:
        vals, f = ctx.assemble(u=problem.u)
    rowptr, colind = ctx.csr_pattern()
    import scipy.sparse as sp
    problem.assembly.K = sp.csr_matrix((vals, colind, rowptr), shape=(n, n))
    if kg is not None:
        problem.assembly.Kg = sp.csr_matrix((kg, colind, rowptr), shape=(n, n))
    problem.assembly.f = f
    return None


def _assemble_dirichlet(problem: Problem, time: float):
    """Nodal collocation (src/problems_dirichlet.jl:60-89): C1 = C2 = I and
    g = value on every DOF that has a prescribed component."""
    rows, vals = [], []
    for s in problem.elements:
        nodes = s.conn.ravel()
        for c in range(3):
            key = f"displacement_{c + 1}"
            if key in s.fields:
                v = np.broadcast_to(np.asarray(s.fields[key], float), nodes.shape)
                rows.append(3 * nodes + c)
                vals.append(v)
    if rows:
        rows = np.concatenate(rows)
        vals = np.concatenate(vals)
    else:
        rows, vals = np.zeros(0, np.int64), np.zeros(0)
    problem.assembly.fixed = (rows, vals)
    return None


class LinearSolver:
    """Solver{Linear} with field and boundary problems (src/solvers.jl)."""

    def __init__(self, name: str = "solver", *problems: Problem):
        self.name = name
        self.problems = list(problems)
        self.u = None
        self.la = None
        self.stats = None
        self.options = dict(method="pcg", precond="block_jacobi", op="matfree", rtol=1e-8)

    def field_problems(self):
        return [p for p in self.problems if isinstance(p.properties, Elasticity)]

    def boundary_problems(self):
        return [p for p in self.problems if isinstance(p.properties, Dirichlet)]


def solve(solver: LinearSolver, **options):
    """solve!(solver) -> (u, la) with u on every DOF and la the reactions on
    the prescribed DOFs (src/solvers.jl:192-216, 273-344).  Options:
    method = "pcg" | "mixed", precond = "block_jacobi" | "learned",
    op = "matfree" | "csr" | "ebe", rtol (relative FP64 residual)."""
    opts = dict(solver.options)
    opts.update(options)
    fields = solver.field_problems()
    if len(fields) != 1:
        raise RuntimeError("the GPU path solves one elasticity field problem")
    body = fields[0]
    if body.assembly.isempty():
        assemble(body, 0.0)
    n = 3 * body.X.shape[0]
    fixed = np.zeros(n, np.uint8)
    g = np.zeros(n)
    for bc in solver.boundary_problems():
        if bc.assembly.isempty():
            assemble(bc, 0.0)
        rows, vals = bc.assembly.fixed
        fixed[rows] = 1
        g[rows] = vals
    if not fixed.any():
        warnings.warn("No rows in C2, forget to set Dirichlet boundary conditions to model?")
    ctx = body.context()
    _, _, _, _, faces, ftype, tr = _elasticity_arrays(body)
    ctx.upload_bcs(fixed, g, faces, ftype, tr)
    method = {"pcg": _lib.SOLVE_PCG, "mixed": _lib.SOLVE_MIXED_IR}[opts["method"]]
    prec = {"block_jacobi": _lib.PREC_BLOCK_JACOBI, "learned": _lib.PREC_LEARNED}[opts["precond"]]
    op, op_outer = opts["op"], opts.get("op_outer", opts["op"])
    if body.properties.finite_strain:
        # K = Km + Kg lives in the context's CSR operator (src/solvers.jl:286-292)
        op = op_outer = "csr"
    try:
        u, la, st = ctx.solve(f=body.assembly.f, method=method, op=op, op_outer=op_outer,
                              precond=prec, rtol=opts["rtol"], inner_rtol=opts.get("inner_rtol", 1e-4),
                              want_la=True)
    except _lib.JfemError as e:
        raise RuntimeError("Failed to solve linear system!") from e
    solver.u, solver.la, solver.stats = u, la, st
    return u, la


def run(solver: LinearSolver, time: float = 0.0, **options):
    """run!(analysis) for a linear static analysis: assemble every problem,
    solve, and update the displacement field (src/solvers.jl:641-657)."""
    for p in solver.problems:
        p.assembly.empty()
        assemble(p, time)
    u, la = solve(solver, **options)
    for p in solver.field_problems():
        p.u = u.copy()
    return u, la


class NonlinearSolver(LinearSolver):
    """Solver{Nonlinear} (src/solvers.jl:536-552): min/max iterations and the
    convergence tolerance on the norm of the solution change."""

    def __init__(self, name: str = "solver", *problems: Problem):
        super().__init__(name, *problems)
        self.time = 0.0
        self.iteration = 0
        self.min_iterations = 1
        self.max_iterations = 10
        self.convergence_tolerance = 5.0e-5
        self.error_if_no_convergence = True
        self.u_norm_change = None


def run_nonlinear(solver: NonlinearSolver, time: float = 0.0, **options):
    """run!(::Solver{Nonlinear}) (src/solvers.jl:575-620): per iteration empty
    and re-assemble every problem, solve the linearised system, add the
    increment to the displacement (incremental formulation,
    src/assembly/problems.jl:288-310) and stop when |du| falls below the
    tolerance (has_converged, :558-572).  Returns True on convergence."""
    for solver.iteration in range(1, solver.max_iterations + 1):
        for p in solver.problems:
            p.assembly.empty()
            assemble(p, time)
        du, la = solve(solver, **options)
        change = 0.0
        for p in solver.field_problems():
            prev = np.zeros_like(du) if p.u is None else p.u
            p.u = prev + du
            change = max(change, float(np.linalg.norm(du)))
            trivial = np.isclose(np.linalg.norm(p.u), 0.0)
        solver.u_norm_change = change
        if solver.iteration >= solver.min_iterations and (change < solver.convergence_tolerance or trivial):
            return True
    if solver.error_if_no_convergence:
        raise RuntimeError(f"nonlinear iteration did not converge in {solver.iteration} iterations!")
    return False
