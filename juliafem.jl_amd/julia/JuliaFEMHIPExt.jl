# JuliaFEMHIPExt.jl -- thin ccall glue between JuliaFEM and libjfem_hip.so.
#
# Overrides the two boundary methods of the elasticity path:
#   assemble_elements!(problem::Problem{Elasticity}, assembly, elements, time)
#       (src/assembly/assembly.jl:25-29 -> src/problems_elasticity.jl:65-81)
#   solve!(solver, K, C1, C2, D, f, g, u, la, ::Type{Val{1}})
#       (src/solvers.jl:192-216)
# All logic lives in the C library (include/jfem.h); this file only converts
# Julia's element/field structures to flat arrays and back.  Julia is not
# available where this package is built, so this file is untested here; the
# same C ABI is exercised by the Python tests.
module JuliaFEMHIPExt

using JuliaFEM
using JuliaFEM: Problem, Elasticity, Assembly, Element, get_connectivity, SparseMatrixCOO
using SparseArrays

const LIB = joinpath(@__DIR__, "..", "libjfem_hip.so")

# field order and types mirror jfm_solve_opts / jfm_newton_opts (include/jfem.h);
# isbits structs get the C layout, padding included
struct SolveOpts
    method::Int32; op::Int32; op_outer::Int32; precond::Int32
    rtol::Float64; inner_rtol::Float64
    maxit::Int32; max_outer::Int32; check_every::Int32; learned_probe::Int32
    inner_maxit::Int32; first_rtol::Float64; coarse_nodes::Int32
end

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

mutable struct SolveStats
    iterations::Int32; outer::Int32; converged::Int32; fallback::Int32
    relres::Float64; solve_ms::Float64; setup_ms::Float64
end
SolveStats() = SolveStats(0, 0, 0, 0, 0.0, 0.0, 0.0)

const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function ctx()
    if CTX[] == C_NULL
        out = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:jfm_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, out)
        rc == 0 || error("jfm_create failed ($rc)")
        CTX[] = out[]
    end
    return CTX[]
end

check(rc) = rc == 0 || error(unsafe_string(ccall((:jfm_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx())))

const NPE = Dict(:Tet10 => 10, :Tet4 => 4, :Hex8 => 8)

# Flatten an element set into (X 3xN, conn npe x n_el, E, nu); node ids are
# kept as given (1-based) so DOF = 3(node-1)+c as in get_gdofs.
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
        E[k] = el("youngs modulus", (0.25, 0.25, 0.25), time)
        nu[k] = el("poissons ratio", (0.25, 0.25, 0.25), time)
    end
    X = zeros(3, nmax)
    for el in elements
        Xe = el("geometry", time)
        for (i, node) in enumerate(get_connectivity(el))
            X[:, node] .= Xe[i]
        end
    end
    return X, conn, E, nu
end

# assemble!(assembly, problem::Problem{Elasticity}, elements, time) for volume
# elements: K and f_ext - f_int computed on the GPU, appended as COO triplets
# of the assembled CSR (sparse() of them reproduces it exactly).
function JuliaFEM.assemble_elements!(problem::Problem{Elasticity}, assembly::Assembly,
                                     elements::Vector{Element{E}}, time) where E
    etype = Symbol(nameof(E))
    haskey(NPE, etype) || return invoke(JuliaFEM.assemble_elements!,
        Tuple{Problem, Assembly, Vector{Element{E}}, Any}, problem, assembly, elements, time)
    X, conn, Ev, nu = flatten(elements, time)
    GC.@preserve X conn Ev nu begin
        check(ccall((:jfm_upload_mesh, LIB), Cint,
              (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Int32),
              ctx(), size(X, 2), X, size(conn, 2), NPE[etype], conn, Ev, nu, 1))
    end
    n = 3 * size(X, 2)
    nnz = Ref{Int64}(0)
    check(ccall((:jfm_csr_pattern, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Int32),
                ctx(), nnz, C_NULL, C_NULL, 1))
    rowptr = Vector{Int64}(undef, n + 1)
    colind = Vector{Int64}(undef, nnz[])
    vals = Vector{Float64}(undef, nnz[])
    f = Vector{Float64}(undef, n)
    check(ccall((:jfm_csr_pattern, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Int32),
                ctx(), nnz, rowptr, colind, 1))
    check(ccall((:jfm_assemble, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx(), C_NULL, vals, f))
    # K symmetric: CSR arrays are the CSC arrays of the same matrix
    K = SparseMatrixCSC(n, n, rowptr, colind, vals)
    I, J, V = findnz(K)
    append!(assembly.K.I, I); append!(assembly.K.J, J); append!(assembly.K.V, V)
    append!(assembly.f.I, 1:n); append!(assembly.f.J, ones(Int, n)); append!(assembly.f.V, f)
    return nothing
end

# solve!(solver, K, C1, C2, D, f, g, u, la, Val{1}) on the GPU: the
# Dirichlet-eliminated system by preconditioned CG instead of ldlt.
function JuliaFEM.solve!(solver::JuliaFEM.Solver, K, C1, C2, D, f, g, u, la, ::Type{Val{1}};
                         method=1, rtol=1e-8)
    nnz(D) == 0 || return false
    n = size(K, 1)
    fixed = zeros(UInt8, n)
    fixed[JuliaFEM.get_nonzero_rows(C2)] .= 0x01
    gv = Vector(g[:, 1])
    fv = Vector(f[:, 1])
    check(ccall((:jfm_upload_bcs, LIB), Cint,
          (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Float64}, Int64, Int32, Ptr{Int64}, Ptr{Float64}, Int32),
          ctx(), fixed, gv, 0, 6, C_NULL, C_NULL, 1))
    opts = Ref(SolveOpts(method, 2, 2, 0, rtol, 1e-4, 100000, 20, 32, 0, 0, 0.0, 512))
    st = Ref(SolveStats())
    rc = ccall((:jfm_solve, LIB), Cint,
               (Ptr{Cvoid}, Ref{SolveOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{SolveStats}),
               ctx(), opts, fv, u, la, st)
    rc == 0 || return false
    @info("GPU solve: $(st[].iterations) iterations, relative residual $(st[].relres)")
    return true
end

# solve_backend!(data, physics; newton_tol, max_newton, ...) for the
# neo-Hookean model (ext/JuliaFEMCUDAExt.jl:685-854, 886-916): the mesh and
# boundary conditions are uploaded by the calls above; returns the reference's
# (u, newton_iters, cg_iters, residual, history) tuple, history holding
# (cumulative CG iterations, relative residual, 0.0) per Newton iterate.
function solve_nonlinear_hip(n::Integer; load_steps=5, newton_tol=1e-8, max_newton=20, method=1, precond=0)
    opts = Ref(NewtonOpts(load_steps, max_newton, newton_tol, method, precond, 0.1, 20000, 0, 1e-2, 0, 32, 4096))
    st = Ref(NewtonStats())
    u = Vector{Float64}(undef, n)
    rc = ccall((:jfm_solve_nonlinear, LIB), Cint,
               (Ptr{Cvoid}, Ref{NewtonOpts}, Ptr{Float64}, Ref{NewtonStats}), ctx(), opts, u, st)
    rc == 0 || @warn("Newton did not converge: " * unsafe_string(ccall((:jfm_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx())))
    m = 1024
    its = Vector{Int32}(undef, m)
    res = Vector{Float64}(undef, m)
    k = ccall((:jfm_solve_history, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Float64}, Int32), ctx(), its, res, m)
    history = [(Int(its[i]), res[i], 0.0) for i in 1:max(k, 0)]
    return (u, Int(st[].newton_iterations), Int(st[].linear_iterations), st[].relres, history)
end

end # module
