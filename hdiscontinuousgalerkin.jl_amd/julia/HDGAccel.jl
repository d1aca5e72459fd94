"""
    HDGAccel

Drop-in GPU backend for the HDG driver of HDiscontinuousGalerkin.jl
(examples/poisson2D_HDG.jl).  The per-cell loop of `doassemble`, `apply!` +
`K \\ b`, `get_uσ!` and `errornorm` run in libhdg (HIP, MI355X) through
`ccall`; meshes and trial functions are the reference's own objects and are
passed without copies (Vector{Node{2,Float64}}, Vector{Cell{2,3,3}} and
PolygonalMesh.faces have exactly the layouts libhdg's C ABI expects).

Usage, replacing poisson2D_HDG.jl:189-218:

    using HDiscontinuousGalerkin, HDGAccel
    ctx = HDGAccel.Context()
    K, b, K_e, b_e = HDGAccel.doassemble(ctx, Vh, Wh, Mh)      # same tuple as doassemble
    apply!(K, b, dbc); û = K \\ b                               # unchanged, or:
    HDGAccel.hdg_solve!(ctx, σ_h, u_h, û_h, dbc)                # fused, device-resident
    Etu_h = HDGAccel.errornorm(ctx)                             # squared L2, as errornorm

Set ENV["LIBHDG"] to the path of libhdg.so before loading.
"""
module HDGAccel

using SparseArrays
import LinearAlgebra
using HDiscontinuousGalerkin

const libhdg = get(ENV, "LIBHDG", "libhdg.so")

struct Options
    k::Int32
    quad_degree::Int32
    tau::Float64
    precision::Int32
    rtol::Float64
    maxit::Int32
    schedule::Int32
    precond::Int32
end

struct Source
    kind::Int32
    kappa_on::Int32
    ka::Float64
    kb::Float64
    coef::NTuple{16,Float64}
    f_qp::Ptr{Float64}
end

struct Sizes
    n_nodes::Int64; n_cells::Int64; n_faces::Int64; n_dofs::Int64; nnz::Int64
    nu::Int32; nq::Int32; nt::Int32; m::Int32; ntr::Int32; nqp::Int32; nfq::Int32
    sym_blocks::Int32
end

mutable struct SolveInfo
    iters::Int32; refinements::Int32; relres::Float64; true_relres::Float64; seconds::Float64
end
SolveInfo() = SolveInfo(0, 0, 0.0, 0.0, 0.0)

mutable struct Context
    h::Ptr{Cvoid}
    function Context(device::Integer = 0, hbm_budget::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:hdg_create, libhdg), Cint, (Ptr{Ptr{Cvoid}}, Cint, Csize_t), r, device, hbm_budget)
        st == 0 || error("hdg_create failed with status $st")
        ctx = new(r[])
        finalizer(c -> ccall((:hdg_destroy, libhdg), Cvoid, (Ptr{Cvoid},), c.h), ctx)
        ctx
    end
end

# status codes -> the exceptions the reference throws
function check(ctx::Context, st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:hdg_last_error, libhdg), Cstring, (Ptr{Cvoid},), ctx.h))
    st in (1, 2, 3, 10) && throw(ArgumentError(msg))        # boundary.jl, ScalarFunctionSpaces.jl:110, quadrature.jl:24
    st == 4 && throw(LinearAlgebra.SingularException(0))   # singular local matrix
    error(msg)
end

"""Uploads a PolygonalMesh{2,3,3} (src/mesh.jl:43-49) without copying it."""
function set_mesh!(ctx::Context, mesh)
    faces = mesh.faces                       # Matrix{Int}: column-major nf x 4
    GC.@preserve mesh faces begin
        check(ctx, ccall((:hdg_set_mesh, libhdg), Cint,
                         (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Int64}),
                         ctx.h, length(mesh.nodes), Ptr{Float64}(pointer(mesh.nodes)),
                         length(mesh.cells), Ptr{Int64}(pointer(mesh.cells)), size(faces, 1), pointer(faces)))
    end
end

order(Wh) = getorder(Wh.fe.func_basis)
quad_degree(Wh) = length(Wh.qr_face_weigths)   # DefaultQuad: GL with quad_degree points

function setup!(ctx::Context, Wh; tau = 1.0, precision = 0, rtol = 1e-12, maxit = 0)
    o = Ref(Options(order(Wh), quad_degree(Wh), tau, precision, rtol, maxit, 0, 0))
    check(ctx, ccall((:hdg_setup, libhdg), Cint, (Ptr{Cvoid}, Ptr{Options}), ctx.h, o))
end

"""f evaluated at every cell quadrature point on the host (function_value,
DiscreteFunctions.jl:6-9), handed to libhdg as HDG_SRC_VALUES."""
function set_source!(ctx::Context, Wh, f)
    mesh = getmesh(Wh)
    nq = getnquadpoints(Wh)
    fq = Matrix{Float64}(undef, getncells(mesh), nq)
    for c in 1:getncells(mesh), q in 1:nq
        fq[c, q] = function_value(f, Wh, c, q)
    end
    GC.@preserve fq begin
        s = Ref(Source(3, 0, 0.0, 0.0, ntuple(_ -> 0.0, 16), pointer(fq)))
        check(ctx, ccall((:hdg_set_source, libhdg), Cint, (Ptr{Cvoid}, Ptr{Source}), ctx.h, s))
    end
end

function sizes(ctx::Context)
    s = Ref{Sizes}()
    check(ctx, ccall((:hdg_query_sizes, libhdg), Cint, (Ptr{Cvoid}, Ptr{Sizes}), ctx.h, s))
    s[]
end

"""doassemble(Vh, Wh, Mh, τ) (poisson2D_HDG.jl:58-186): returns
(K::SparseMatrixCSC, rhs, K_element, b_element) computed on the GPU."""
function doassemble(ctx::Context, Vh, Wh, Mh, τ = 1.0; f = x -> 2 * π^2 * sin(π * x[1]) * sin(π * x[2]))
    set_mesh!(ctx, getmesh(Wh))
    setup!(ctx, Wh; tau = τ)
    set_source!(ctx, Wh, f)
    check(ctx, ccall((:hdg_assemble_condense, libhdg), Cint, (Ptr{Cvoid},), ctx.h))
    s = sizes(ctx)
    colptr = Vector{Int64}(undef, s.n_dofs + 1)
    rowval = Vector{Int64}(undef, s.nnz)
    nzval = Vector{Float64}(undef, s.nnz)
    rhs = Vector{Float64}(undef, s.n_dofs)
    check(ctx, ccall((:hdg_get_system, libhdg), Cint,
                     (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, colptr, rowval, nzval, rhs))
    Ke = Array{Float64}(undef, s.m, s.ntr, s.n_cells)
    be = Array{Float64}(undef, s.m, s.n_cells)
    check(ctx, ccall((:hdg_get_local_ops, libhdg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, Ke, be))
    K = SparseMatrixCSC(s.n_dofs, s.n_dofs, colptr, rowval, nzval)
    K, rhs, [Ke[:, :, c] for c in 1:s.n_cells], [be[:, c] for c in 1:s.n_cells]
end

function set_dirichlet!(ctx::Context, dbc)
    check(ctx, ccall((:hdg_set_dirichlet, libhdg), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Float64}),
                     ctx.h, length(dbc.prescribed_dofs), dbc.prescribed_dofs, dbc.values))
end

"""Dirichlet(û_h, mesh, "boundary", g) (boundary.jl:11-42) computed on the
GPU for g(x) = c0 + c1 x + c2 y (g_kind 0, par = (c0, c1, c2)) or
sin(πx) sin(πy) (g_kind 1); reference = true reproduces the reference's own
weights and accumulator instead of the projection."""
function set_dirichlet_function!(ctx::Context, g_kind::Integer, par::NTuple{3,Float64}; reference::Bool = false)
    p = collect(par)
    check(ctx, ccall((:hdg_set_dirichlet_function, libhdg), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint),
                     ctx.h, g_kind, p, reference ? 1 : 0))
end

"""apply!(K, b, dbc) (boundary.jl:121-175) done on the GPU to the assembled
system: returns (K, b) with the Dirichlet rows and columns zeroed,
meandiag(K) on their diagonal and b lifted, as a SparseMatrixCSC and a
vector (the device's own system is left unchanged)."""
function applied_system(ctx::Context)
    s = sizes(ctx)
    colptr = Vector{Int64}(undef, s.n_dofs + 1)
    rowval = Vector{Int64}(undef, s.nnz)
    nzval = Vector{Float64}(undef, s.nnz)
    rhs = Vector{Float64}(undef, s.n_dofs)
    md = Ref(0.0)
    check(ctx, ccall((:hdg_get_applied_system, libhdg), Cint,
                     (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, colptr, rowval, nzval, rhs, md))
    SparseMatrixCSC(s.n_dofs, s.n_dofs, colptr, rowval, nzval), rhs
end

function recover!(ctx::Context, σ_h, u_h, û_h)
    check(ctx, ccall((:hdg_recover, libhdg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, σ_h.m_values, u_h.m_values, û_h.m_values))
end

"""get_uσ!(σ_h, u_h, û_h, û, ...) (poisson2D_HDG.jl:197-212) on the GPU for a
trace solution û computed by the caller."""
function get_uσ!(ctx::Context, σ_h, u_h, û_h, û::Vector{Float64})
    check(ctx, ccall((:hdg_set_trace, libhdg), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.h, û))
    recover!(ctx, σ_h, u_h, û_h)
end

"""Fused assemble + condense + Dirichlet-eliminated PCG + recovery; fills the
trial functions and returns the solve record."""
function hdg_solve!(ctx::Context, σ_h, u_h, û_h, dbc; τ = 1.0, f = x -> 2 * π^2 * sin(π * x[1]) * sin(π * x[2]),
                    precision = 0, rtol = 1e-12)
    Wh = u_h.fs
    set_mesh!(ctx, getmesh(Wh))
    setup!(ctx, Wh; tau = τ, precision = precision, rtol = rtol)
    set_source!(ctx, Wh, f)
    set_dirichlet!(ctx, dbc)
    info = Ref(SolveInfo())
    check(ctx, ccall((:hdg_run, libhdg), Cint, (Ptr{Cvoid}, Ptr{SolveInfo}), ctx.h, info))
    recover!(ctx, σ_h, u_h, û_h)
    info[]
end

"""errornorm(u_h, u_ex) for u_ex = sin(πx)sin(πy): the squared L2 error
(DiscreteFunctions.jl:97-120) of the last recovered u_h, on the GPU."""
function errornorm(ctx::Context)
    out = Ref(0.0)
    check(ctx, ccall((:hdg_errornorm, libhdg), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, 0, C_NULL, out))
    out[]
end

# ---- batches (hdg_batch_*; no counterpart in the reference, which solves one
# problem per script run).  Request and Result mirror hdg_request / hdg_result.

struct Request
    nx::Int64; ny::Int64; jitter::Float64; seed::UInt64
    opt::Options; src::Source; uex_kind::Int32; want_trace::Int32
end

struct Result
    id::Int64; status::Int32; iters::Int32
    relres::Float64; true_relres::Float64; errornorm::Float64
    n_cells::Int64; n_dofs::Int64
    t_submit::Float64; t_start::Float64; t_done::Float64
    device_bytes::Csize_t; slot::Int32
end

"""
    Batch(reqs; device = 0, hbm_budget = 0)

Many Poisson solves on rectangle meshes in grouped launches: requests of one
order and options are packed into one mesh and share every kernel launch;
each result is bitwise that of the request solved alone.
"""
mutable struct Batch
    h::Ptr{Cvoid}
    n::Int
    function Batch(reqs::Vector{Request}; device::Integer = 0, hbm_budget::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:hdg_batch_create, libhdg), Cint, (Ptr{Ptr{Cvoid}}, Cint, Csize_t, Cint, Ptr{Request}),
                   r, device, hbm_budget, length(reqs), reqs)
        if st != 0
            msg = r[] == C_NULL ? "" :
                  unsafe_string(ccall((:hdg_batch_last_error, libhdg), Cstring, (Ptr{Cvoid},), r[]))
            r[] == C_NULL || ccall((:hdg_batch_destroy, libhdg), Cvoid, (Ptr{Cvoid},), r[])
            error("hdg_batch_create failed with status $st: $msg")
        end
        b = new(r[], length(reqs))
        finalizer(x -> ccall((:hdg_batch_destroy, libhdg), Cvoid, (Ptr{Cvoid},), x.h), b)
        b
    end
end

"""Condense, assemble, solve and recover every instance; one Result per request."""
function run!(b::Batch)
    res = Vector{Result}(undef, b.n)
    st = ccall((:hdg_batch_run, libhdg), Cint, (Ptr{Cvoid}, Ptr{Result}), b.h, res)
    if st != 0
        msg = unsafe_string(ccall((:hdg_batch_last_error, libhdg), Cstring, (Ptr{Cvoid},), b.h))
        @warn "hdg_batch_run: status $st ($msg); see each Result.status"
    end
    res
end

"""Trace solution of request i (1-based) after run!."""
function trace(b::Batch, i::Integer, n_dofs::Integer)
    out = Vector{Float64}(undef, n_dofs)
    st = ccall((:hdg_batch_trace, libhdg), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64), b.h, i - 1, out, length(out))
    st == 0 || error("hdg_batch_trace failed with status $st")
    out
end

end # module
