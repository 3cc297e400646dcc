"""
    MadSDPHIP

Julia binding of the gfx950 library (`libmadsdp_hip.so`, C ABI in
`include/madsdp_hip.h`) as a MadSDP Schur-solver component.  Selecting
`schur = MadSDPHIP.HIPSchur()` in `MadSDPSolver(model; ...)` (or through JuMP
with `set_attribute(model, "schur", MadSDPHIP.HIPSchur())`) runs the whole
interior-point loop of that solver on the GPU: the problem is read once through
LowRankOpt's accessors, uploaded, solved on the device with the same
direction / step / centering / initial-point components, and the iterate, the
scalars and the status are written back into the Julia `MadSDPSolver`, so
`MadSDP.Optimizer`, the MOI front end and `solver.stats` are unchanged.

The entry points mirror the reference's driver: `solve!` (src/solver.jl:385),
`initialize!` (src/solver.jl:286), `iterate!` (src/primal_dual.jl:85,104;
src/dual_centering.jl:266) and `update_termination!` (src/primal_dual.jl:123;
src/dual_centering.jl:293) are specialised on solvers whose Schur workspace is
a `HIPWorkspace`.
"""
module MadSDPHIP

import LinearAlgebra
import SparseArrays
import LowRankOpt as LRO
import MadNLP
import MadSDP

const libmsdp = get(ENV, "MADSDP_HIP_LIB",
                    joinpath(@__DIR__, "..", "madsdp_amd", "libmadsdp_hip.so"))

# --- C structs (same layout as include/madsdp_hip.h) -------------------------
mutable struct Options
    tol::Cdouble; max_iter::Int32; max_wall_time::Cdouble; mu_min::Cdouble
    regularize_schur::Cdouble; max_schur_reg::Int32
    algorithm::Int32; direction::Int32; step::Int32; centering::Int32
    initial_point::Int32; schur::Int32
    tau::Cdouble; shrink::Cdouble; boundary::Cdouble; sigma_exponent::Cdouble
    initial_ratio::Cdouble; ratio::Cdouble; identity_scale::Cdouble
    initial_shift::Cdouble; max_shift::Cdouble
    dc_beta::Cdouble; dc_delta::Cdouble; dc_A::Cdouble; dc_centering_max_iter::Int32
    precision::Int32; mu_switch::Cdouble; refine_max_iter::Int32
    Options() = new()
end

mutable struct Info
    mu::Cdouble; sigma::Cdouble; alpha::Cdouble; beta::Cdouble
    inf_pr::Cdouble; inf_du::Cdouble; inf_compl::Cdouble; obj_val::Cdouble
    pobj::Cdouble; dobj::Cdouble; iter::Int32; status::Int32
    n_newton::Int32; n_predictor::Int32; dc_t::Cdouble; dc_lambda::Cdouble
    lowp_factorizations::Int32; fp64_fallbacks::Int32; refine_steps::Int32; pad::Int32
    Info() = new()
end

const MSDP_X, MSDP_S, MSDP_Y, MSDP_XLIN, MSDP_ZLIN = 0, 1, 2, 3, 4

function check(rc, what)
    if rc < 0
        msg = unsafe_string(ccall((:msdp_last_error, libmsdp), Cstring, ()))
        error("MadSDPHIP $what: $msg")
    end
    return rc
end

# --- the component ------------------------------------------------------------
"""
    HIPSchur(; device = 0, precision = :fp64)

Runs the solver on HIP device `device`.  `precision = :mixed` assembles the
dense constraint blocks' Schur contributions in FP32 on MFMA while
μ ≥ `mu_switch` and refines every Schur solve against the FP64 operator.
"""
struct HIPSchur <: MadSDP.AbstractSchurSolver
    device::Int
    precision::Symbol
end
HIPSchur(; device = 0, precision = :fp64) = HIPSchur(device, precision)

mutable struct HIPWorkspace{T}
    handle::Ptr{Cvoid}
    device::Int
    precision::Int32
end

function MadSDP.schur_workspace(s::HIPSchur, ::Type{T}, m) where {T}
    T == Float64 || throw(ArgumentError("HIPSchur requires Float64 data."))
    ws = HIPWorkspace{T}(C_NULL, s.device, s.precision == :mixed ? Int32(1) : Int32(0))
    finalizer(ws) do w
        w.handle == C_NULL || ccall((:msdp_solver_destroy, libmsdp), Cint, (Ptr{Cvoid},), w.handle)
    end
    return ws
end
MadSDP.check_schur(::HIPSchur, ::MadSDP.AbstractAlgorithm) = nothing

const HIPSolver = MadSDP.MadSDPSolver{T,M,B,A,AS,I,<:HIPWorkspace} where {T,M,B,A,AS,I}

# component -> C ABI selectors (include/madsdp_hip.h)
function options(solver::HIPSolver)
    o = Options()
    ccall((:msdp_default_options, libmsdp), Cvoid, (Ref{Options},), o)
    opt = solver.opt
    o.tol, o.max_iter, o.max_wall_time = opt.tol, opt.max_iter, opt.max_wall_time
    o.mu_min, o.regularize_schur, o.max_schur_reg = opt.mu_min, opt.regularize_schur, opt.max_schur_reg
    a = solver.algorithm
    if a isa MadSDP.DualCentering
        o.algorithm = 1
        o.dc_beta, o.dc_delta = a.β, a.δ
        o.dc_A = a.A === nothing ? -1.0 : a.A
        o.dc_centering_max_iter = a.centering_max_iter
    else
        o.algorithm = 0
        o.direction = a.direction isa MadSDP.NT ? 0 : 1
        if a.step isa MadSDP.EigenStep
            o.step, o.tau = 0, a.step.τ
        else
            o.step, o.shrink, o.boundary = 1, a.step.shrink, a.step.boundary
        end
        if a.centering isa MadSDP.Mehrotra
            o.centering, o.sigma_exponent = 0, a.centering.σ_exponent
        else
            o.centering, o.initial_ratio, o.ratio = 1, a.centering.initial_ratio, a.centering.ratio
        end
    end
    p = solver.initial_point
    if p isa MadSDP.ScaledIdentity
        o.initial_point = 0
        o.identity_scale = p.scale === nothing ? 0.0 : p.scale
    elseif p isa MadSDP.InitialPoint
        o.initial_point = 1
    else
        o.initial_point, o.initial_shift, o.max_shift = 2, p.initial_shift, p.max_shift
    end
    o.schur = 0
    o.precision = solver.schur.precision
    return o
end

# LowRankOpt.Model -> device upload (the accessors src/qr_schur.jl:24-37 uses)
function upload!(solver::HIPSolver)
    ws = solver.schur
    model = MadSDP._inner(solver.model)
    dims = Int64[b.n for b in solver.blocks]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:msdp_solver_create, libmsdp), Cint,
                (Ref{Ptr{Cvoid}}, Cint, Int32, Ptr{Int64}, Int64, Int64),
                h, ws.device, length(dims), dims, solver.m, solver.n_scalar), "create")
    ws.handle = h[]
    for (i, idx) in enumerate(LRO.matrix_indices(model))
        C = Matrix{Float64}(LRO.grad(model, idx))
        check(ccall((:msdp_solver_set_cost_block, libmsdp), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                    ws.handle, i - 1, C), "set_cost_block")
        con, row, col, val = Int64[], Int64[], Int64[], Float64[]
        for j = 1:solver.m
            A = LRO.jac(model, j, idx)
            iszero(A) && continue
            I, J, V = SparseArrays.findnz(SparseArrays.sparse(A))
            append!(con, fill(j - 1, length(V))); append!(row, I .- 1)
            append!(col, J .- 1); append!(val, V)
        end
        check(ccall((:msdp_solver_set_constraints_coo, libmsdp), Cint,
                    (Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                    ws.handle, i - 1, length(val), con, row, col, val), "set_constraints_coo")
    end
    b = Vector{Float64}(LRO.cons_constant(solver.model))
    check(ccall((:msdp_solver_set_rhs, libmsdp), Cint, (Ptr{Cvoid}, Ptr{Float64}), ws.handle, b), "set_rhs")
    if solver.n_scalar > 0
        d = Vector{Float64}(model.d_lin)
        Cl = Matrix{Float64}(model.C_lin)      # m x n_scalar, column-major
        check(ccall((:msdp_solver_set_scalar_data, libmsdp), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ws.handle, d, Cl), "set_scalar_data")
    end
    o = options(solver)
    check(ccall((:msdp_solver_configure, libmsdp), Cint, (Ptr{Cvoid}, Ref{Options}), ws.handle, o),
          "configure")
    p = solver.initial_point
    if p isa MadSDP.InitialPoint
        X = reduce(vcat, [vec(Matrix{Float64}(Xi)) for Xi in p.X])
        x = p.x === nothing ? C_NULL : Vector{Float64}(p.x)
        check(ccall((:msdp_solver_set_initial_point, libmsdp), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    ws.handle, X, Vector{Float64}(p.y), x), "set_initial_point")
    end
    return ws.handle
end

# device iterate and scalars -> the Julia solver (what primal_objective,
# print_iter and MOI.get read)
function sync_back!(solver::HIPSolver)
    h = solver.schur.handle
    for (i, b) in enumerate(solver.blocks)
        check(ccall((:msdp_solver_get, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                    h, MSDP_X, i - 1, b.X), "get X")
        check(ccall((:msdp_solver_get, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                    h, MSDP_S, i - 1, b.S), "get S")
    end
    check(ccall((:msdp_solver_get, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                h, MSDP_Y, -1, solver.y), "get y")
    if solver.n_scalar > 0
        check(ccall((:msdp_solver_get, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                    h, MSDP_XLIN, -1, solver.x), "get x")
        check(ccall((:msdp_solver_get, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                    h, MSDP_ZLIN, -1, solver.z), "get z")
    end
    inf = Info()
    check(ccall((:msdp_solver_get_info, libmsdp), Cint, (Ptr{Cvoid}, Ref{Info}), h, inf), "get_info")
    solver.mu, solver.sigma, solver.alpha, solver.beta = inf.mu, inf.sigma, inf.alpha, inf.beta
    solver.inf_pr, solver.inf_du, solver.inf_compl = inf.inf_pr, inf.inf_du, inf.inf_compl
    solver.obj_val, solver.iter = inf.obj_val, inf.iter
    solver.status = MadNLP.Status(inf.status)
    return solver
end

handle(solver::HIPSolver) =
    solver.schur.handle == C_NULL ? upload!(solver) : solver.schur.handle

function MadSDP.initialize!(solver::HIPSolver)
    check(ccall((:msdp_solver_initialize, libmsdp), Cint, (Ptr{Cvoid},), handle(solver)), "initialize!")
    return sync_back!(solver)
end

function MadSDP.update_termination!(solver::HIPSolver)
    check(ccall((:msdp_solver_update_termination, libmsdp), Cint, (Ptr{Cvoid},), handle(solver)),
          "update_termination!")
    return sync_back!(solver)
end

for A in (MadSDP.PrimalDual, MadSDP.DualCentering)
    @eval function MadSDP.iterate!(solver::HIPSolver, ::$A, args...)
        check(ccall((:msdp_solver_iterate, libmsdp), Cint, (Ptr{Cvoid},), handle(solver)), "iterate!")
        return sync_back!(solver)
    end
end

function MadSDP.solve!(solver::HIPSolver; verbose::Bool = true)
    check(ccall((:msdp_solver_solve, libmsdp), Cint, (Ptr{Cvoid},), handle(solver)), "solve!")
    sync_back!(solver)
    verbose && MadSDP.print_iter(solver)
    return solver
end

end # module
