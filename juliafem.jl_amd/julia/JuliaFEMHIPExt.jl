# JuliaFEMHIPExt.jl -- ccall glue between JuliaFEM and libjfem_hip.so.
#
# Overrides the boundary methods of the elasticity path:
#   assemble_elements!(problem::Problem{Elasticity}, assembly, elements, time)
#       (src/assembly/assembly.jl:25-29 -> src/problems_elasticity.jl:65-81, 203-451)
#   solve!(solver, K, C1, C2, D, f, g, u, la, ::Type{Val{1}})
#       (src/solvers.jl:192-216)
#   initialize_backend(::GPU, physics, time), solve_backend!(data, physics; ...)
#       (src/backend/abstract.jl:181-241; ext/JuliaFEMCUDAExt.jl:867-916)
# All numerics live in the C library (include/jfem.h).  Every override hands
# the work back to the reference method (invoke, or `false` from solve!) when
# the case is outside what the library computes exactly as the reference
# does; the conditions are listed at each method.  Julia is not installed
# where this package is built: tests/test_gpu_julia_glue.py replays this
# file's ccall sequence through ctypes against the oracle.
module JuliaFEMHIPExt

using JuliaFEM
using JuliaFEM: Problem, Elasticity, Assembly, Element, get_connectivity, get_integration_points
using JuliaFEM: GPU, Physics, ElasticityPhysicsType, AbstractElasticityData
using SparseArrays
using LinearAlgebra
using Random

const LIB = joinpath(@__DIR__, "..", "libjfem_hip.so")

const JFM_OP_CSR = Int32(0)
const JFM_OP_MATFREE = Int32(2)
const JFM_FP64 = Int32(0)

# field order and types mirror jfm_solve_opts / jfm_solve_stats (include/jfem.h);
# isbits structs get the C layout, padding included
struct SolveOpts
    method::Int32; op::Int32; op_outer::Int32; precond::Int32
    rtol::Float64; inner_rtol::Float64
    maxit::Int32; max_outer::Int32; check_every::Int32; learned_probe::Int32
    inner_maxit::Int32; first_rtol::Float64; coarse_nodes::Int32
end

mutable struct SolveStats
    iterations::Int32; outer::Int32; converged::Int32; fallback::Int32
    relres::Float64; solve_ms::Float64; setup_ms::Float64
end
SolveStats() = SolveStats(0, 0, 0, 0, 0.0, 0.0, 0.0)

# FP32 two-level PCG inside FP64 refinement (the library's fastest exact-to-rtol method)
default_opts(rtol) = SolveOpts(1, JFM_OP_MATFREE, JFM_OP_MATFREE, 0, rtol, 2e-3, 200000, 50, 16, 0, 0, 0.0, 1536)

function new_ctx()
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:jfm_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, out)
    rc == 0 || error("jfm_create failed ($rc)")
    return out[]
end

last_error(c) = unsafe_string(ccall((:jfm_last_error, LIB), Cstring, (Ptr{Cvoid},), c))
check(c, rc) = rc == 0 || error("libjfem_hip: " * last_error(c))

# One library context per (problem, volume element type): each holds that
# group's mesh, so a problem with several element groups assembles each into
# its own context and the COO triplets sum as in the reference.  `tl` records
# whether the last assembly was the finite-strain one (its K lives in the
# context's CSR operator only).
mutable struct Group
    ctx::Ptr{Cvoid}
    tl::Bool
end

const GROUPS = Dict{Tuple{UInt,Symbol},Group}()

function group(problem, etype::Symbol)
    key = (objectid(problem), etype)
    return get!(GROUPS, key) do
        Group(new_ctx(), false)
    end
end

problem_groups(problem) = [g for ((p, _), g) in GROUPS if p == objectid(problem)]

const NPE = Dict(:Tet10 => 10, :Tet4 => 4, :Hex8 => 8)

# A per-element constant material: E and nu at every integration point equal.
function element_material(el, time)
    E = nu = NaN
    for ip in get_integration_points(el)
        Ei = el("youngs modulus", ip, time)
        ni = el("poissons ratio", ip, time)
        if isnan(E)
            E, nu = Ei, ni
        elseif Ei != E || ni != nu
            return nothing
        end
    end
    return (E, nu)
end

# Flatten an element group into (X 3xN, conn npe x n_el, E, nu, u 3xN), node
# ids kept as given (1-based) so DOF = 3(node-1)+c as in get_gdofs
# (src/assembly/problems.jl:466-478); nothing when the material varies
# inside an element.
function flatten(elements, time)
    n_el = length(elements)
    npe = length(first(elements))
    conn = Matrix{Int64}(undef, npe, n_el)
    E = Vector{Float64}(undef, n_el)
    nu = Vector{Float64}(undef, n_el)
    nmax = 0
    for (k, el) in enumerate(elements)
        c = get_connectivity(el)
        conn[:, k] .= c
        nmax = max(nmax, maximum(c))
        mat = element_material(el, time)
        mat === nothing && return nothing
        E[k], nu[k] = mat
    end
    X = zeros(3, nmax)
    U = zeros(3, nmax)
    for el in elements
        Xe = el("geometry", time)
        ue = haskey(el, "displacement") ? el("displacement", time) : nothing
        for (i, node) in enumerate(get_connectivity(el))
            X[:, node] .= Xe[i]
            ue === nothing || (U[:, node] .= ue[i])
        end
    end
    return X, conn, E, nu, U
end

# Cases the library does not compute: they run the reference assembly.
function reference_only(problem, elements, etype)
    props = problem.properties
    haskey(NPE, etype) || return true                       # surface and other element types
    props.formulation == :continuum || return true          # plane stress / plane strain
    props.geometric_stiffness && !props.finite_strain && return true   # Kg of the small strain
    isempty(props.store_fields) || return true              # stress/strain stored at points
    el = first(elements)
    for name in ("plasticity", "displacement load", "displacement load 1",
                 "displacement load 2", "displacement load 3")
        haskey(el, name) && return true
    end
    return false
end

# assemble!(assembly, problem::Problem{Elasticity}, elements, time) for a
# volume element group: K and f = f_ext - f_int(u) (f_ext = 0: volume groups
# carry no load here, see reference_only) computed on the GPU and appended as
# the COO triplets of the assembled CSR; sparse() of them reproduces it exactly
# (src/problems_elasticity.jl:239-242, 432-433; src/sparse/sparse.jl:121-150).
# With finite_strain the material tangent goes to assembly.K and, with
# geometric_stiffness, Kg to assembly.Kg, as the reference does (:441-447).
function JuliaFEM.assemble_elements!(problem::Problem{Elasticity}, assembly::Assembly,
                                     elements::Vector{Element{E}}, time) where E
    etype = Symbol(nameof(E))
    fallback() = invoke(JuliaFEM.assemble_elements!,
                        Tuple{Problem, Assembly, Vector{Element{E}}, Any}, problem, assembly, elements, time)
    reference_only(problem, elements, etype) && return fallback()
    flat = flatten(elements, time)
    flat === nothing && return fallback()
    X, conn, Ev, nu, U = flat
    grp = group(problem, etype)
    c = grp.ctx
    GC.@preserve X conn Ev nu begin
        check(c, ccall((:jfm_upload_mesh, LIB), Cint,
              (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Int32),
              c, size(X, 2), X, size(conn, 2), NPE[etype], conn, Ev, nu, 1))
    end
    n = 3 * size(X, 2)
    nnz_ = Ref{Int64}(0)
    check(c, ccall((:jfm_csr_pattern, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Int32),
                   c, nnz_, C_NULL, C_NULL, 1))
    rowptr = Vector{Int64}(undef, n + 1)
    colind = Vector{Int64}(undef, nnz_[])
    vals = Vector{Float64}(undef, nnz_[])
    f = Vector{Float64}(undef, n)
    check(c, ccall((:jfm_csr_pattern, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Int32),
                   c, nnz_, rowptr, colind, 1))
    uvec = vec(U)      # 3 x N column-major = DOF order 3(node-1)+c
    props = problem.properties
    kg = props.geometric_stiffness ? Vector{Float64}(undef, nnz_[]) : Float64[]
    GC.@preserve uvec begin
        if props.finite_strain
            check(c, ccall((:jfm_assemble_tl, LIB), Cint,
                           (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                           c, Int32(props.geometric_stiffness), uvec, vals,
                           props.geometric_stiffness ? kg : C_NULL, f))
        else
            check(c, ccall((:jfm_assemble, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                           c, uvec, vals, f))
        end
    end
    grp.tl = props.finite_strain
    # K symmetric: the CSR arrays are the CSC arrays of the same matrix
    append_coo!(assembly.K, SparseMatrixCSC(n, n, rowptr, colind, vals))
    props.geometric_stiffness && append_coo!(assembly.Kg, SparseMatrixCSC(n, n, copy(rowptr), copy(colind), kg))
    append!(assembly.f.I, 1:n); append!(assembly.f.J, ones(Int, n)); append!(assembly.f.V, f)
    return nothing
end

function append_coo!(A, K)
    I, J, V = findnz(K)
    append!(A.I, I); append!(A.J, J); append!(A.V, V)
end

# The context whose K is the K handed to solve!: the solver's field problems
# must have exactly one GPU-assembled group, with the same size, and K x must
# equal the device's K x for two random x (K = 1/2 (K + K') + Kg + Kb
# reduces to it only when Kg and Kb vanish).
function matching_group(solver, K)
    gs = Group[]
    for p in JuliaFEM.get_field_problems(solver)
        append!(gs, problem_groups(p))
    end
    length(gs) == 1 || return nothing
    c = gs[1].ctx
    n = size(K, 1)
    info = zeros(Int64, 16)
    check(c, ccall((:jfm_get_info, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}), c, info))
    info[3] == n || return nothing                          # jfm_info.n_dofs
    length(JuliaFEM.get_nonzero_rows(K)) == n || return nothing   # every DOF has stiffness
    rng = MersenneTwister(20)
    for _ in 1:2
        x = randn(rng, n)
        y = Vector{Float64}(undef, n)
        check(c, ccall((:jfm_upload_bcs, LIB), Cint,
              (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Int32),
              c, zeros(UInt8, n), C_NULL, 0, 6, C_NULL, C_NULL, 1))
        check(c, ccall((:jfm_apply, LIB), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}),
                       c, JFM_OP_CSR, JFM_FP64, x, y))
        Kx = K * x
        norm(Kx - y) <= 1e-12 * norm(Kx) || return nothing
    end
    return gs[1]
end

is_identity(A) = size(A, 1) == size(A, 2) && A == sparse(1.0I, size(A, 1), size(A, 1))

# solve!(solver, K, C1, C2, D, f, g, u, la, Val{1}) on the GPU: the same
# elimination as the reference (u[B] = g[B], K[I,I] u[I] = f[I] - K[I,B] u[B],
# la[B] = f[B] - K[B,:] u), solved to a relative residual of `rtol` by the
# library instead of ldlt.  Returns false -- so the reference's ldlt runs --
# when D is not empty, B != B2, C1 or C2 is not the identity on B, or K is not
# the matrix the library assembled (matching_group).  A finite-strain K is
# solved with the CSR operator that holds it, a linear one matrix-free.
function JuliaFEM.solve!(solver::JuliaFEM.Solver, K, C1, C2, D, f, g, u, la, ::Type{Val{1}};
                         rtol=1e-8)
    nnz(D) == 0 || return false
    B = JuliaFEM.get_nonzero_rows(C2)
    B2 = JuliaFEM.get_nonzero_columns(C2)
    B == B2 || return false
    isempty(B) || (is_identity(C2[B, B2]) && is_identity(C1[B2, B])) || return false
    grp = matching_group(solver, K)
    grp === nothing && return false
    c = grp.ctx
    n = size(K, 1)
    fixed = zeros(UInt8, n)
    fixed[B] .= 0x01
    gv = Vector{Float64}(g[:, 1])
    fv = Vector{Float64}(f[:, 1])
    check(c, ccall((:jfm_upload_bcs, LIB), Cint,
          (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Int32),
          c, fixed, gv, 0, 6, C_NULL, C_NULL, 1))
    o = default_opts(rtol)
    op = grp.tl ? JFM_OP_CSR : JFM_OP_MATFREE
    opts = Ref(SolveOpts(o.method, op, op, o.precond, o.rtol, o.inner_rtol, o.maxit, o.max_outer,
                         o.check_every, o.learned_probe, o.inner_maxit, o.first_rtol, o.coarse_nodes))
    st = Ref(SolveStats())
    uu = Vector{Float64}(undef, n)
    ll = Vector{Float64}(undef, n)
    rc = ccall((:jfm_solve, LIB), Cint,
               (Ptr{Cvoid}, Ref{SolveOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{SolveStats}),
               c, opts, fv, uu, ll, st)
    if rc != 0
        @warn("GPU solve failed, using the reference solver: " * last_error(c))
        return false
    end
    u[:] .= uu
    la[B] .= ll[B]
    @info("GPU solve: $(st[].iterations) iterations, relative residual $(st[].relres)")
    return true
end

# ---- new API backend (src/backend/abstract.jl:181-241) -------------------

mutable struct ElasticityDataHIP <: AbstractElasticityData
    ctx::Ptr{Cvoid}
    n_nodes::Int
    node_list::Vector{Int}      # renumbered node k is node_list[k]
end

# initialize_backend(::GPU, physics, time): the mesh, materials, Dirichlet
# values and surface tractions of `physics`, nodes renumbered in ascending id
# order as initialize_gpu_data! does (ext/JuliaFEMCUDAExt.jl:86-217), uploaded
# once.  Tet4, Tet10 and Hex8 bodies; Tri3, Tri6 and Quad4 load faces.
function JuliaFEM.initialize_backend(::GPU, physics::Physics{ElasticityPhysicsType}, time::Float64)
    els = physics.body_elements
    isempty(els) && error("No body elements in physics!")
    npe = length(get_connectivity(first(els)))
    npe in (4, 8, 10) || error("unsupported body element with $npe nodes")
    node_list = sort(unique(vcat([collect(get_connectivity(el)) for el in els]...)))
    node_map = Dict(nd => k for (k, nd) in enumerate(node_list))
    N = length(node_list)
    X = zeros(3, N)
    conn = Matrix{Int64}(undef, npe, length(els))
    E = Vector{Float64}(undef, length(els))
    nu = Vector{Float64}(undef, length(els))
    for (e, el) in enumerate(els)
        Xe = el.fields.geometry
        for (i, nd) in enumerate(get_connectivity(el))
            conn[i, e] = node_map[nd]
            X[:, node_map[nd]] .= Xe[:, i]
        end
        E[e] = el.fields.youngs_modulus
        nu[e] = el.fields.poissons_ratio
    end
    fixed = zeros(UInt8, 3N)
    g = zeros(3N)
    bc = physics.bc_dirichlet
    for (i, nd) in enumerate(bc.node_ids)
        haskey(node_map, nd) || continue
        for (k, comp) in enumerate(bc.components[i])
            dof = 3 * (node_map[nd] - 1) + comp
            fixed[dof] = 0x01
            g[dof] = bc.values[i][k]
        end
    end
    surf = physics.bc_neumann.surface_elements
    nf = length(surf)
    nfn = nf == 0 ? 3 : length(get_connectivity(first(surf)))
    fconn = Matrix{Int64}(undef, nfn, nf)
    tr = zeros(3, nf)
    for (i, s) in enumerate(surf)
        for (j, nd) in enumerate(get_connectivity(s))
            haskey(node_map, nd) || error("Surface element references node $nd not in body mesh")
            fconn[j, i] = node_map[nd]
        end
        tr[:, i] .= physics.bc_neumann.traction[i][1:3]
    end
    c = new_ctx()
    check(c, ccall((:jfm_upload_mesh, LIB), Cint,
          (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Int32),
          c, N, X, length(els), npe, conn, E, nu, 1))
    check(c, ccall((:jfm_upload_bcs, LIB), Cint,
          (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Int32),
          c, fixed, g, nf, nfn, nf == 0 ? C_NULL : fconn, nf == 0 ? C_NULL : tr, 1))
    data = ElasticityDataHIP(c, N, node_list)
    finalizer(d -> ccall((:jfm_destroy, LIB), Cvoid, (Ptr{Cvoid},), d.ctx), data)
    return data
end

# solve_backend!(data, physics; tol, max_iter, newton_tol, max_newton,
# max_cg_per_newton): the linear solve the reference's backend performs
# (its Newton loop stops after the first solve for linear elasticity,
# ext/JuliaFEMCUDAExt.jl:685-854), converged when the absolute residual norm
# |f_ext - K u| on free DOFs is below newton_tol, as there.  Returns the
# reference's tuple (u, newton_iterations, cg_iterations, residual, history).
function JuliaFEM.solve_backend!(data::ElasticityDataHIP, physics::Physics{ElasticityPhysicsType};
                                 tol=1e-6, max_iter=1000, newton_tol=1e-6, max_newton=20, max_cg_per_newton=50)
    c = data.ctx
    n = 3 * data.n_nodes
    f = Vector{Float64}(undef, n)
    check(c, ccall((:jfm_assemble, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   c, C_NULL, C_NULL, f))
    fixed = zeros(UInt8, n)
    bc = physics.bc_dirichlet
    node_map = Dict(nd => k for (k, nd) in enumerate(data.node_list))
    for (i, nd) in enumerate(bc.node_ids)
        haskey(node_map, nd) || continue
        for comp in bc.components[i]
            fixed[3 * (node_map[nd] - 1) + comp] = 0x01
        end
    end
    R0 = norm(f[fixed .== 0x00])
    u = zeros(n)
    R0 == 0.0 && return (u, 1, 0, 0.0, Tuple{Int,Float64,Float64}[])
    rtol = clamp(newton_tol / R0, 1e-13, 0.5)
    opts = Ref(default_opts(rtol))
    st = Ref(SolveStats())
    rc = ccall((:jfm_solve, LIB), Cint,
               (Ptr{Cvoid}, Ref{SolveOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{SolveStats}),
               c, opts, C_NULL, u, C_NULL, st)
    rc == 0 || rc == -4 || check(c, rc)
    residual = st[].relres * R0
    iters = Int(st[].iterations)
    return (u, 1, iters, residual, [(iters, R0, rtol)])
end

# The compressible neo-Hookean Newton of the library (jfm_solve_nonlinear;
# src/materials/neo_hookean.jl:129-231, the Newton-Krylov loop of
# ext/JuliaFEMCUDAExt.jl:685-854 with an FP64 residual) on a backend context:
# load-stepped Newton from u = 0, returning the same 5-tuple as
# solve_backend!, history = (cumulative CG iterations, relative residual, 0.0)
# per Newton iterate.
struct NewtonOpts
    load_steps::Int32; max_newton::Int32; rtol::Float64; method::Int32; precond::Int32
    eta_max::Float64; maxit_linear::Int32; learned_probe::Int32; inner_rtol::Float64
    inner_maxit::Int32; check_every::Int32; coarse_nodes::Int32
end

mutable struct NewtonStats
    newton_iterations::Int32; linear_iterations::Int32; converged::Int32; fallbacks::Int32
    relres::Float64; solve_ms::Float64
end
NewtonStats() = NewtonStats(0, 0, 0, 0, 0.0, 0.0)

function solve_nonlinear_hip(data::ElasticityDataHIP; load_steps=5, newton_tol=1e-8, max_newton=20)
    c = data.ctx
    opts = Ref(NewtonOpts(load_steps, max_newton, newton_tol, 1, 0, 0.1, 20000, 0, 1e-2, 0, 32, 4096))
    st = Ref(NewtonStats())
    u = Vector{Float64}(undef, 3 * data.n_nodes)
    rc = ccall((:jfm_solve_nonlinear, LIB), Cint,
               (Ptr{Cvoid}, Ref{NewtonOpts}, Ptr{Float64}, Ref{NewtonStats}), c, opts, u, st)
    rc == 0 || @warn("Newton did not converge: " * last_error(c))
    m = 1024
    its = Vector{Int32}(undef, m)
    res = Vector{Float64}(undef, m)
    k = ccall((:jfm_solve_history, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Float64}, Int32), c, its, res, m)
    history = [(Int(its[i]), res[i], 0.0) for i in 1:max(k, 0)]
    return (u, Int(st[].newton_iterations), Int(st[].linear_iterations), st[].relres, history)
end

end # module
