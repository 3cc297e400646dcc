"""
    MadSDPHIP

Julia binding of the gfx950 library (`libmadsdp_hip.so`, C ABI in
`include/madsdp_hip.h`) as MadSDP components, plugged in through the
reference's own component seam; `MadSDPSolver`, `solve!`, `iterate!` and the
MOI / LowRankOpt front end are the stock ones.

* `HIPSchur <: MadSDP.AbstractSchurSolver` overrides the two methods a Schur
  solver may take over (src/components.jl:253-266, the pattern of
  src/qr_schur.jl:24,90):
  - `factorize_schur!(ws, solver)` pushes the current iterate to the device,
    forms there the block factorizations and the NT/HKM scaling, assembles the
    Schur complement (W∘W for one-hot constraints, MFMA products for dense
    ones) and factors it with the reference's regularisation ladder
    (src/solver.jl:299-314);
  - `solve_directions!(ws, solver, μ, corrected)` runs rhs!, the Schur solve
    and directions! on the device (the Mehrotra correction there too when
    `corrected`) and fills `solver.dely`, `b.work.delX`, `b.work.delS`,
    `solver.delx`, `solver.delz`, as the stock driver expects;
  - `schur_solve!(ws, rhs)` (src/schur_solvers.jl:21), which DualCentering
    calls directly (src/dual_centering.jl:107,126,239).
* `HIPNT <: MadSDP.AbstractDirection` and `HIPEigenStep <: MadSDP.AbstractStepLength`
  (optional, with `HIPSchur`) leave the NT scaling and the step-length
  eigenvalues to the device: their per-block host methods do nothing and
  `step_lengths` reads the device's (α, β).

Usage: `MadSDPSolver(model; schur = MadSDPHIP.HIPSchur())`, or through JuMP
`set_attribute(model, "schur", MadSDPHIP.HIPSchur())`, optionally with
`direction = MadSDPHIP.HIPNT(), step = MadSDPHIP.HIPEigenStep()`.
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
    lowp_min_gflop::Cdouble
    shared_device::Int32
    Options() = new()
end

# iterate arrays of msdp_solver_get / msdp_solver_set
const MSDP_X, MSDP_S, MSDP_Y, MSDP_XLIN, MSDP_ZLIN = 0, 1, 2, 3, 4
const MSDP_DELX, MSDP_DELS, MSDP_DELY, MSDP_DELXLIN, MSDP_DELZLIN = 5, 6, 7, 8, 9

function check(rc, what)
    if rc < 0
        msg = unsafe_string(ccall((:msdp_last_error, libmsdp), Cstring, ()))
        error("MadSDPHIP $what: $msg")
    end
    return rc
end

getarr!(h, which, block, dst) =
    check(ccall((:msdp_solver_get, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                h, which, block, dst), "get $which")
setarr!(h, which, block, src) =
    check(ccall((:msdp_solver_set, libmsdp), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                h, which, block, src), "set $which")

# --- components ---------------------------------------------------------------
"""
    HIPSchur(; device = 0, precision = :fp64)

Schur complement on HIP device `device`.  `precision = :mixed` assembles large
dense constraint blocks in FP32 on MFMA while μ ≥ `mu_switch` and refines
every Schur solve against the FP64 operator; `precision = :bf16x3` forms the
same blocks' H = B'B from bf16 hi/lo splits of the FP64 scaled panel B and
refines through B.
"""
struct HIPSchur <: MadSDP.AbstractSchurSolver
    device::Int
    precision::Symbol
end
HIPSchur(; device = 0, precision = :fp64) = HIPSchur(device, precision)

"NT scaling computed on the device by `HIPSchur` (requires `schur = HIPSchur()`)."
struct HIPNT <: MadSDP.AbstractDirection end
"EigenStep(τ) evaluated on the device by `HIPSchur` (requires `schur = HIPSchur()`)."
struct HIPEigenStep{T} <: MadSDP.AbstractStepLength
    τ::T
end
HIPEigenStep(; τ = 0.95) = HIPEigenStep(τ)

mutable struct HIPWorkspace{T}
    handle::Ptr{Cvoid}
    device::Int
    precision::Int32
end

function MadSDP.schur_workspace(s::HIPSchur, ::Type{T}, m) where {T}
    T == Float64 || throw(ArgumentError("HIPSchur requires Float64 data."))
    prec = s.precision == :mixed ? Int32(1) : s.precision == :bf16x3 ? Int32(2) : Int32(0)
    ws = HIPWorkspace{T}(C_NULL, s.device, prec)
    finalizer(ws) do w
        w.handle == C_NULL || ccall((:msdp_solver_destroy, libmsdp), Cint, (Ptr{Cvoid},), w.handle)
    end
    return ws
end

MadSDP.check_schur(::HIPSchur, ::MadSDP.AbstractAlgorithm) = nothing

# HIPNT / HIPEigenStep: the device does the per-block work inside
# factorize_schur!; the host methods the stock driver calls per block are no-ops
MadSDP.direction_workspace(::HIPNT, ::Type{T}, n) where {T} = HIPNT()
MadSDP.prepare_direction!(::HIPNT, ::MadSDP.SDPBlock) = true
MadSDP.schur_scaling(::HIPNT, b::MadSDP.SDPBlock) = (b.Si, b.Si)   # unused by HIPSchur
MadSDP.correction!(corr, ::HIPNT, ::MadSDP.SDPBlock) = corr          # formed on the device
MadSDP.step_workspace(::HIPEigenStep, ::Type{T}, n) where {T} = nothing
MadSDP.prepare_step!(::HIPEigenStep, ::MadSDP.SDPBlock) = true
function MadSDP.step_lengths(::HIPEigenStep, solver)
    α, β = Ref(0.0), Ref(0.0)
    check(ccall((:msdp_solver_step_lengths, libmsdp), Cint, (Ptr{Cvoid}, Ref{Cdouble}, Ref{Cdouble}),
                solver.schur.handle, α, β), "step_lengths")
    return α[], β[]
end

# component -> C ABI selectors (include/madsdp_hip.h)
function options(solver, ws::HIPWorkspace)
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
        o.direction = a.direction isa MadSDP.HKM ? 1 : 0
        if a.step isa MadSDP.CholeskyBacktrack
            o.step, o.shrink, o.boundary = 1, a.step.shrink, a.step.boundary
        else
            o.step, o.tau = 0, a.step.τ
        end
        if a.centering isa MadSDP.Mehrotra
            o.centering, o.sigma_exponent = 0, a.centering.σ_exponent
        else
            o.centering, o.initial_ratio, o.ratio = 1, a.centering.initial_ratio, a.centering.ratio
        end
    end
    o.schur = 0
    o.precision = ws.precision
    return o
end

# LowRankOpt keeps factored constraint matrices as factor * Diagonal(scaling) *
# factor' (its rank-one detection, README.md:148): (factors n x r, weights r),
# or nothing for any other matrix type.
function _factors(A)
    (hasproperty(A, :factor) && hasproperty(A, :scaling)) || return nothing
    F = Matrix{Float64}(reshape(A.factor, size(A, 1), :))
    return F, Vector{Float64}(A.scaling)
end

# LowRankOpt.Model -> device, once (the accessors src/qr_schur.jl:25-37 uses).
# A block whose constraint matrices are all factored goes up as factors;
# otherwise one whose matrices have more than `datasparsity` nonzeros
# (src/options.jl:12-14) goes up dense: its m_i touching matrices as one
# contiguous m_i x n x n array; otherwise as COO entries of both triangles.
function upload!(ws::HIPWorkspace, solver)
    model = MadSDP._inner(solver.model)
    dims = Int64[b.n for b in solver.blocks]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:msdp_solver_create, libmsdp), Cint,
                (Ref{Ptr{Cvoid}}, Cint, Int32, Ptr{Int64}, Int64, Int64),
                h, ws.device, length(dims), dims, solver.m, solver.n_scalar), "create")
    ws.handle = h[]
    κ = solver.opt.datasparsity
    for (i, idx) in enumerate(LRO.matrix_indices(model))
        C = Matrix{Float64}(LRO.grad(model, idx))
        check(ccall((:msdp_solver_set_cost_block, libmsdp), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                    ws.handle, i - 1, C), "set_cost_block")
        n = dims[i]
        touching = [(j, LRO.jac(model, j, idx)) for j = 1:solver.m]
        filter!(t -> !iszero(t[2]), touching)
        facs = [_factors(A) for (_, A) in touching]
        if !isempty(facs) && all(!isnothing, facs)
            con = Int64[j - 1 for (j, _) in touching]
            ptr = cumsum(Int64[0; [length(w) for (_, w) in facs]])
            F = reduce(hcat, [f for (f, _) in facs])     # n x R, column-major
            w = reduce(vcat, [w for (_, w) in facs])
            check(ccall((:msdp_solver_set_constraints_lowrank, libmsdp), Cint,
                        (Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
                        ws.handle, i - 1, length(con), con, ptr, F, w), "set_constraints_lowrank")
            continue
        end
        maxnnz = maximum((SparseArrays.nnz(SparseArrays.sparse(A)) for (_, A) in touching); init = 0)
        if maxnnz > κ
            con = Int64[j - 1 for (j, _) in touching]
            mats = Array{Float64}(undef, n, n, length(touching))
            for (k, (_, A)) in enumerate(touching)
                mats[:, :, k] .= A
            end
            check(ccall((:msdp_solver_set_constraints_dense, libmsdp), Cint,
                        (Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Float64}),
                        ws.handle, i - 1, length(con), con, mats), "set_constraints_dense")
        else
            con, row, col, val = Int64[], Int64[], Int64[], Float64[]
            for (j, A) in touching
                I, J, V = SparseArrays.findnz(SparseArrays.sparse(A))
                append!(con, fill(j - 1, length(V))); append!(row, I .- 1)
                append!(col, J .- 1); append!(val, V)
            end
            check(ccall((:msdp_solver_set_constraints_coo, libmsdp), Cint,
                        (Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                        ws.handle, i - 1, length(val), con, row, col, val), "set_constraints_coo")
        end
    end
    b = Vector{Float64}(LRO.cons_constant(solver.model))
    check(ccall((:msdp_solver_set_rhs, libmsdp), Cint, (Ptr{Cvoid}, Ptr{Float64}), ws.handle, b), "set_rhs")
    if solver.n_scalar > 0
        d = Vector{Float64}(model.d_lin)
        Cl = Matrix{Float64}(model.C_lin)      # m x n_scalar, column-major
        check(ccall((:msdp_solver_set_scalar_data, libmsdp), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ws.handle, d, Cl), "set_scalar_data")
    end
    check(ccall((:msdp_solver_configure, libmsdp), Cint, (Ptr{Cvoid}, Ref{Options}), ws.handle,
                options(solver, ws)), "configure")
    check(ccall((:msdp_solver_upload, libmsdp), Cint, (Ptr{Cvoid},), ws.handle), "upload")
    return ws.handle
end

# the host iterate (after the stock _take_step!, src/primal_dual.jl:56-72) -> device
function push_iterate!(ws::HIPWorkspace, solver)
    h = ws.handle
    for (i, b) in enumerate(solver.blocks)
        setarr!(h, MSDP_X, i - 1, b.X)
        setarr!(h, MSDP_S, i - 1, b.S)
    end
    setarr!(h, MSDP_Y, -1, solver.y)
    if solver.n_scalar > 0
        setarr!(h, MSDP_XLIN, -1, solver.x)
        setarr!(h, MSDP_ZLIN, -1, solver.z)
    end
    check(ccall((:msdp_solver_update_residuals, libmsdp), Cint, (Ptr{Cvoid},), h), "update_residuals")
    return
end

# factorize_schur! (src/solver.jl:295-314; override as src/qr_schur.jl:24)
function MadSDP.factorize_schur!(ws::HIPWorkspace, solver)
    ws.handle == C_NULL && upload!(ws, solver)
    push_iterate!(ws, solver)
    ok = Ref{Int32}(0)
    check(ccall((:msdp_solver_prepare, libmsdp), Cint, (Ptr{Cvoid}, Ref{Int32}), ws.handle, ok),
          "prepare")
    ok[] == 0 && return false
    check(ccall((:msdp_solver_factorize_schur, libmsdp), Cint, (Ptr{Cvoid}, Ref{Int32}), ws.handle, ok),
          "factorize_schur")
    return ok[] != 0
end

# solve_directions! (src/primal_dual.jl:44-54; override as src/qr_schur.jl:90)
function MadSDP.solve_directions!(ws::HIPWorkspace, solver, μ, corrected::Bool)
    h = ws.handle
    # the Mehrotra correction of the predictor directions still on the device
    # (src/step.jl:128-136, src/directions.jl:90-100)
    corrected && check(ccall((:msdp_solver_correction, libmsdp), Cint, (Ptr{Cvoid},), h), "correction")
    check(ccall((:msdp_solver_solve_directions, libmsdp), Cint, (Ptr{Cvoid}, Cdouble, Int32),
                h, μ, corrected ? 1 : 0), "solve_directions")
    getarr!(h, MSDP_DELY, -1, solver.dely)
    for (i, b) in enumerate(solver.blocks)
        getarr!(h, MSDP_DELX, i - 1, b.work.delX)
        getarr!(h, MSDP_DELS, i - 1, b.work.delS)
    end
    if solver.n_scalar > 0
        getarr!(h, MSDP_DELXLIN, -1, solver.delx)
        getarr!(h, MSDP_DELZLIN, -1, solver.delz)
    end
    return
end

# schur_solve! (src/schur_solvers.jl:21) on the last factorization
function MadSDP.schur_solve!(ws::HIPWorkspace, rhs)
    v = rhs isa Vector{Float64} ? rhs : Vector{Float64}(rhs)
    check(ccall((:msdp_solver_schur_solve, libmsdp), Cint, (Ptr{Cvoid}, Ptr{Float64}), ws.handle, v),
          "schur_solve")
    v === rhs || copyto!(rhs, v)
    return rhs
end

end # module
