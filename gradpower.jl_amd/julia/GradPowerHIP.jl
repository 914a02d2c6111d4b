# GradPowerHIP.jl -- Julia front end of the MI355X batched backward-Euler
# integrator (libgradpower_hip.so, C ABI in include/gradpower_hip.h).
#
# Drop-in counterparts of the GPU entry points of GradPower.jl's CUDA
# extension (ext/GradPowerCUDAExt/GradPowerCUDAExt.jl):
#   GpuBatchedLayout(dp, ps, M)                       ext:466        -> HipBatchedLayout(dp, ps, M)
#   integrate_gpu_cudss!(gbl, ps, tf; dt, newton_tol) ext:2924-2996  -> integrate_hip!(hbl, ps, tf; ...)
#   integrate_gpu_schur_cudss!(gbl, ps, tf; ...)      ext:3008-3080  -> integrate_hip_schur!(hbl, ps, tf; fp32, ...)
#   integrate_gpu_shared_multi!(gbl, ps, tf, fault_buses, rfaults, ton, toff; ...)
#                                                     ext:2640-2708  -> integrate_hip_multi!(hbl, ps, tf, ...)
#   GradPower.adjoint(λ0, dprob, sys, traj, tvec; functional=true)
#                  (examples/adjoint.jl:34-51,106)                  -> hip_adjoint, hip_objective + rrule
# Return shapes follow the reference: (tvec, trajs) with trajs[m] an
# n_sys x (nsteps+1) matrix in the cluster-reordered z order.
#
# Usage (GradPower.jl set up as usual: from_psse, build_network!, runpf!,
# DynamicProblem, initialize_dynamics!):
#   include("GradPowerHIP.jl"); using .GradPowerHIP
#   hbl = HipBatchedLayout(dp, ps, 256)
#   tvec, trajs = integrate_hip_schur!(hbl, ps, 5.0; fp32 = true)
module GradPowerHIP

using GradPower
using SparseArrays
using ChainRulesCore

export HipBatchedLayout, integrate_hip!, integrate_hip_schur!, integrate_hip_multi!, integrate_hip_shared!, integrate_hip_shared_multi!, hip_adjoint,
       hip_objective, scenario_status, symbolic_info

const LIB = joinpath(@__DIR__, "..", "gradpower_hip", "libgradpower_hip.so")

# ---- C structs: field order and types exactly as in include/gradpower_hip.h ----
struct GpSystemDesc
    n::Int32
    nbus::Int32
    net_ptr::Int32
    n_par::Int32
    is_diff::Ptr{UInt8}
    y_nnz::Int32
    y_rowptr::Ptr{Int32}
    y_colind::Ptr{Int32}
    y_val::Ptr{Float64}
    n_dev::NTuple{8,Int32}
    dev::NTuple{8,Ptr{Int32}}
    genrou_u0::Ptr{Float64}
    n_clusters::Int32
    clusters::Ptr{Int32}
end

struct GpEvent
    scenario::Int32
    bus::Int32
    r_fault::Float64
    step_on::Int32
    step_off::Int32
end

struct GpLineTrip
    scenario::Int32
    from_bus::Int32
    to_bus::Int32
    step_trip::Int32
    step_reclose::Int32
    reserved::Int32
    y::NTuple{8,Float64}          # yff, yft, ytf, ytt as (re, im)
end

struct GpLoadDrop
    scenario::Int32
    load::Int32
    step::Int32
    reserved::Int32
end

struct GpDeviceDrop
    scenario::Int32
    type::Int32              # GP_ZIPLOAD, GP_SEXS or GP_ESDC1A
    row::Int32
    step::Int32
end

struct GpOptions
    record_all::Int32
    n_record::Int32
    record_idx::Ptr{Int32}
    refactor_policy::Int32
    stall_window::Int32
    profile::Int32
    lock_step::Int32
    dense_block::Int32
    retry_window::Int32
    global_x::Int32
    refactor_theta::Float64
    async_below::Float64
    predictor::Int32
    pivot_rule::Int32
    solve_levels::Int32
end

struct GpStats
    newton_iters::Int64
    scenario_steps::Int64
    event_resolves::Int64
    refactorizations::Int64
    n_not_converged::Int32
    n_nan::Int32
    max_final_residual::Float64
    elapsed_ms::Float64
    solve_ms::Float64
    newton_rounds::Int64
    factor_launches::Int64
    solve_kernel_ms::Float64
    solve_kernel_launches::Int64
    solve_kernel_scenarios::Int64
    factor_kernel_ms::Float64
    factor_kernel_launches::Int64
    factor_kernel_scenarios::Int64
    solve_bytes_per_scenario::Float64
    factor_bytes_per_scenario::Float64
    exact_retries::Int64
    repivots::Int64
end

struct GpObjective
    n_rows::Int32
    rows::Ptr{Int32}
    scale::Float64
    offset::Float64
end

struct GpAdjointStats
    elapsed_ms::Float64
    forward_ms::Float64
    backward_ms::Float64
    scenario_steps::Int64
    adjoint_steps::Int64
    n_segments::Int32
    n_excluded::Int32
    newton_iters::Int64
    reserved::NTuple{4,Int64}
end

struct GpSymbolicInfo
    n::Int32
    nnz_j::Int32
    nnz_lu_j::Int32
    levels_j::Int32
    n_red::Int32
    nnz_s::Int32
    nnz_lu_s::Int32
    levels_s::Int32
    n_clusters_nt::Int32
    w_total_nt::Int32
    nd_j::Int32
    nd_s::Int32
    nd_jt::Int32
    res_j::Int32
    res_s::Int32
    ld_scen::Int32
    solve_levels_j::Int32
    solve_levels_s::Int32
end

const GP_OK = Cint(0)
const GP_ERR_NEWTON_NOT_CONVERGED = Cint(-5)
const METHOD_MONO_FP64, METHOD_SCHUR_FP64, METHOD_SCHUR_FP32, METHOD_SCHUR_SHARED = Int32(0), Int32(1), Int32(2), Int32(3)
const REFACTOR_EVERY_ITER, REFACTOR_FIRST_ITER, REFACTOR_ADAPTIVE, REFACTOR_ASYNC = Int32.(0:3)

last_error() = unsafe_string(ccall((:gp_last_error, LIB), Cstring, ()))
check(rc) = rc == GP_OK || error("gradpower_hip error $rc: ", last_error())

# ---- system tables --------------------------------------------------------------
# The 0-based SoA tables of include/gradpower_hip.h from a set-up, reordered
# system (set_dynamics!, then reorder_state!, src/clusters.jl:210-320): after the
# reorder, dev.diff_ptr is an absolute 1-based z index, dev.alg_ptr counts from
# diff_dim (src/clusters.jl:292-301), dev.ctrl_ptr indexes uvec_idx, whose entries
# are z indices (0: unwired, src/GradPower.jl:400, src/coupling.jl:154-256), and
# dev.par_ptr indexes pvec.
struct Tables
    n::Int32
    nbus::Int32
    net_ptr::Int32
    n_par::Int32
    is_diff::Vector{UInt8}
    yrp::Vector{Int32}
    yci::Vector{Int32}
    yv::Vector{Float64}
    dev::NTuple{8,Vector{Int32}}     # row-major [n_t x cols_t]
    n_dev::NTuple{8,Int32}
    u0::Vector{Float64}
    clusters::Vector{Int32}          # row-major [n_clusters x 11]
    n_clusters::Int32
end

const GP_GENROU, GP_IEESGO, GP_TGOV1, GP_SEXS, GP_ESDC1A, GP_IEEEST, GP_ZIPLOAD, GP_STATICGEN = 0:7
const COLS = (6, 4, 4, 4, 4, 4, 2, 4)

devtype(d) = d isa GradPower.Genrou ? GP_GENROU :
             d isa GradPower.IEESGO ? GP_IEESGO :
             d isa GradPower.TGOV1 ? GP_TGOV1 :
             d isa GradPower.SEXS ? GP_SEXS :
             d isa GradPower.ESDC1A ? GP_ESDC1A :
             d isa GradPower.IEEEST ? GP_IEEEST :
             d isa GradPower.ZIPLoad ? GP_ZIPLOAD :
             d isa GradPower.StaticGenerator ? GP_STATICGEN :
             error("GradPowerHIP: device type $(typeof(d)) is not supported")

function build_tables(dp::GradPower.DynamicProblem, ps::GradPower.PowerSystem)
    psd = ps.dynamic
    dd = psd.diff_dim
    nda = psd.diff_dim + psd.alg_dim
    nbus = length(ps.buses)
    n = nda + 2 * nbus
    npt = nda                                     # 0-based z index of vr of bus 1
    is_diff = zeros(UInt8, n)
    for i in psd.diff_indices
        is_diff[i] = 0x01
    end
    # realified Y (2nbus x 2nbus, src/network.jl:87) in CSR: the CSC of its transpose
    Yt = sparse(transpose(ps.network.ybus_real))
    yrp = Int32.(Yt.colptr .- 1)
    yci = Int32.(Yt.rowval .- 1)
    yv = Vector{Float64}(Yt.nzval)
    zdiff(dev) = dev.dtype.diff_size > 0 ? Int32(dev.diff_ptr - 1) : Int32(-1)
    zalg(dev) = dev.dtype.alg_size > 0 ? Int32(dd + dev.alg_ptr - 1) : Int32(-1)
    uidx(j) = (j < 1 || j > length(psd.uvec_idx) || psd.uvec_idx[j] == 0) ? Int32(-1) :
              Int32(psd.uvec_idx[j] - 1)
    rows = ntuple(_ -> Vector{Vector{Int32}}(), 8)
    localrow = zeros(Int, length(psd.devices))
    u0 = Float64[]
    for (i, dev) in enumerate(psd.devices)
        t = devtype(dev.dtype)
        b = Int32(psd.map.bus[i] - 1)
        par = Int32(dev.par_ptr - 1)
        localrow[i] = length(rows[t+1])              # 0-based row in its table
        r = if t == GP_GENROU
            # controls: e_fd at ctrl_ptr, p_m at ctrl_ptr + 1 (src/kernels/genrou.jl:242-254)
            push!(u0, dp.uvec[dev.ctrl_ptr], dp.uvec[dev.ctrl_ptr+1])
            Int32[zdiff(dev), zalg(dev), par, b, uidx(dev.ctrl_ptr + 1), uidx(dev.ctrl_ptr)]
        elseif t in (GP_IEESGO, GP_TGOV1, GP_IEEEST)
            Int32[zdiff(dev), zalg(dev), par, uidx(dev.ctrl_ptr)]     # the machine speed
        elseif t in (GP_SEXS, GP_ESDC1A)
            Int32[zdiff(dev), par, Int32(npt + 2 * b), Int32(-1)]     # vs wired below
        elseif t == GP_ZIPLOAD
            Int32[b, par]
        else
            Int32[b, zalg(dev), par, Int32(dev.dtype.bus_type)]
        end
        push!(rows[t+1], r)
    end
    # cluster table (src/clusters.jl:16-28), and each stabiliser's output into
    # its exciter's vs column (src/tables/ieeest.jl:95-130)
    cl = Int32[]
    ncl = 0
    for c in psd.clusters.clusters
        gt = devtype(psd.devices[c.gen_idx].dtype)
        r = Int32[c.w_start - 1, c.w_size, c.bus - 1, c.trivial ? 1 : 0, -1, 0, -1, 0, -1, -1, -1]
        if gt == GP_GENROU
            r[5] = localrow[c.gen_idx]
        else
            r[11] = localrow[c.gen_idx]
        end
        if c.gov_idx > 0
            r[6] = devtype(psd.devices[c.gov_idx].dtype)
            r[7] = localrow[c.gov_idx]
        end
        if c.exc_idx > 0
            et = devtype(psd.devices[c.exc_idx].dtype)
            r[8] = et
            r[9] = localrow[c.exc_idx]
            if c.pss_idx > 0
                rows[et+1][localrow[c.exc_idx]+1][4] = zalg(psd.devices[c.pss_idx])
            end
        end
        c.pss_idx > 0 && (r[10] = localrow[c.pss_idx])
        append!(cl, r)
        ncl += 1
    end
    flat(v, k) = isempty(v) ? zeros(Int32, 1) : reduce(vcat, v)
    dev = ntuple(t -> flat(rows[t], COLS[t]), 8)
    n_dev = ntuple(t -> Int32(length(rows[t])), 8)
    return Tables(n, nbus, npt, length(dp.pvec), is_diff, yrp, yci, yv, dev, n_dev,
                  isempty(u0) ? zeros(2) : u0, cl, ncl)
end

# ---- layout ------------------------------------------------------------------
mutable struct HipBatchedLayout
    ctx::Ptr{Cvoid}
    M::Int
    n::Int
    zvec0::Vector{Float64}
    pvec::Vector{Float64}
    nrec::Int
    last_stats::Union{Nothing,GpStats}
end

"""
    HipBatchedLayout(dp, ps, M; refactor_policy = REFACTOR_ASYNC, record_idx = nothing)

`GpuBatchedLayout(dp, ps, M)` (ext:466) on the HIP library: M scenarios of one
system, every one starting from `dp.zvec`.  `record_idx` (1-based z indices)
limits the recorded trajectory channels; `nothing` records every state.
"""
function HipBatchedLayout(dp::GradPower.DynamicProblem, ps::GradPower.PowerSystem, M::Int;
                          refactor_policy::Integer = REFACTOR_ASYNC, record_idx = nothing,
                          retry_window::Integer = 0, predictor::Integer = 0)
    t = build_tables(dp, ps)
    rec = record_idx === nothing ? Int32[] : Int32.(collect(record_idx) .- 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve t rec begin
        desc = GpSystemDesc(t.n, t.nbus, t.net_ptr, t.n_par, pointer(t.is_diff), Int32(length(t.yci)),
                            pointer(t.yrp), pointer(t.yci), pointer(t.yv), t.n_dev,
                            ntuple(i -> pointer(t.dev[i]), 8), pointer(t.u0), t.n_clusters, pointer(t.clusters))
        opt = GpOptions(record_idx === nothing ? 1 : 0, length(rec), isempty(rec) ? C_NULL : pointer(rec),
                        Int32(refactor_policy), 0, 0, 0, 0, Int32(retry_window), Int32(0), 0.0, 0.0,
                        Int32(predictor), Int32(0), Int32(0))
        check(ccall((:gp_create, LIB), Cint, (Ref{GpSystemDesc}, Int32, Ref{GpOptions}, Ref{Ptr{Cvoid}}),
                    desc, Int32(M), opt, h))
    end
    hbl = HipBatchedLayout(h[], M, Int(t.n), Vector{Float64}(dp.zvec), Vector{Float64}(dp.pvec),
                           record_idx === nothing ? Int(t.n) : length(rec), nothing)
    finalizer(x -> x.ctx == C_NULL || ccall((:gp_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.ctx), hbl)
    set_state!(hbl, hbl.pvec)
    return hbl
end

set_state!(hbl, p) = check(ccall((:gp_set_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}),
                                 hbl.ctx, hbl.zvec0, Int32(0), p))

steps(t, dt) = Int(round(t / dt))        # round half to even, as src/dynamics.jl:516-542

# line trips and ZIP-load disconnects of ps.dynamic in every scenario (the
# reference runs them in its CPU integrate!, src/dynamics.jl:533-540)
function _topology_events!(hbl, ps, dt)
    psd = ps.dynamic
    tr = [GpLineTrip(Int32(m - 1), Int32(t.from_bus - 1), Int32(t.to_bus - 1), Int32(steps(t.ton, dt)),
                     Int32(isfinite(t.toff) ? steps(t.toff, dt) : 0), Int32(0),
                     (reim(t.yff)..., reim(t.yft)..., reim(t.ytf)..., reim(t.ytt)...))
          for m in 1:hbl.M for t in psd.trip_events]
    isempty(tr) || check(ccall((:gp_set_line_trips, LIB), Cint, (Ptr{Cvoid}, Ptr{GpLineTrip}, Int32),
                               hbl.ctx, tr, Int32(length(tr))))
    # DisconnectDeviceEvent: ZIP loads and SEXS / ESDC1A exciters (the
    # library refuses devices with algebraic states)
    function target(i)
        T = typeof(psd.devices[i].dtype)
        ty = T <: GradPower.ZIPLoad ? GP_ZIPLOAD : T <: GradPower.SEXS ? GP_SEXS :
             T <: GradPower.ESDC1A ? GP_ESDC1A : error("a $(T) cannot be disconnected (algebraic states)")
        return Int32(ty), Int32(count(d -> typeof(d.dtype) == T, psd.devices[1:i-1]))
    end
    dr = [GpDeviceDrop(Int32(m - 1), target(d.device_idx)..., Int32(steps(d.ton, dt)))
          for m in 1:hbl.M for d in psd.disconnect_events]
    isempty(dr) || check(ccall((:gp_set_device_drops, LIB), Cint, (Ptr{Cvoid}, Ptr{GpDeviceDrop}, Int32),
                               hbl.ctx, dr, Int32(length(dr))))
end

function _run!(hbl, ps, method, ev::Vector{GpEvent}, nsteps::Vector{Int32}, dt, newton_tol, itermax)
    set_state!(hbl, hbl.pvec)
    check(ccall((:gp_set_events, LIB), Cint, (Ptr{Cvoid}, Ptr{GpEvent}, Int32, Ptr{Int32}),
                hbl.ctx, ev, Int32(length(ev)), nsteps))
    _topology_events!(hbl, ps, dt)
    st = Ref{GpStats}()
    rc = ccall((:gp_integrate, LIB), Cint, (Ptr{Cvoid}, Int32, Float64, Float64, Int32, Ref{GpStats}),
               hbl.ctx, method, dt, newton_tol, Int32(itermax), st)
    # a scenario whose Newton solve failed stops (gp_get_scenario_status says
    # where); the others are complete
    rc == GP_OK || rc == GP_ERR_NEWTON_NOT_CONVERGED || check(rc)
    hbl.last_stats = st[]
    return [begin
                ln = Int(nsteps[m]) + 1
                out = Matrix{Float64}(undef, ln, hbl.nrec)      # out[k, j]: channel j at step k-1
                check(ccall((:gp_get_trajectory, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64),
                            hbl.ctx, Int32(m - 1), out, Int64(ln)))
                permutedims(out)                                 # n x (nsteps+1), as ext:3075-3079
            end for m in 1:hbl.M]
end

function _shared_events(hbl, ps, dt)
    [GpEvent(Int32(m - 1), Int32(e.bus - 1), e.rfault, Int32(steps(e.ton, dt)), Int32(steps(e.toff, dt)))
     for m in 1:hbl.M for e in ps.dynamic.events]
end

"""
    integrate_hip!(hbl, ps, tf; dt = 1/120, newton_tol = 1e-10, itermax = 30)

`integrate_gpu_cudss!` (ext:2924-2996): plain FP64 monolithic Newton, sparse
LU of J refactorised at every iteration, the events of `ps.dynamic` in every
scenario.  Returns `(tvec, trajs)`.
"""
function integrate_hip!(hbl::HipBatchedLayout, ps, tf::Float64; dt::Float64 = 1.0 / 120.0,
                        newton_tol::Float64 = 1e-10, itermax::Int = 30)
    ns = steps(tf, dt)
    trajs = _run!(hbl, ps, METHOD_MONO_FP64, _shared_events(hbl, ps, dt), fill(Int32(ns), hbl.M), dt,
                  newton_tol, itermax)
    for e in ps.dynamic.events
        GradPower.deactivate!(e)      # as ext:3071-3073
    end
    return collect(0:dt:tf), trajs
end

"""
    integrate_hip_schur!(hbl, ps, tf; fp32 = false, ...)

`integrate_gpu_schur_cudss!` (ext:3008-3080): cluster Schur reduction; with
`fp32 = true` the block and reduced-network factors are FP32 while residuals,
updates and states stay FP64 (inexact Newton; a failed solve is run again as
exact FP64 Newton before its scenario counts as failed).
"""
function integrate_hip_schur!(hbl::HipBatchedLayout, ps, tf::Float64; dt::Float64 = 1.0 / 120.0,
                              newton_tol::Float64 = 1e-10, itermax::Int = 30, fp32::Bool = false)
    ns = steps(tf, dt)
    trajs = _run!(hbl, ps, fp32 ? METHOD_SCHUR_FP32 : METHOD_SCHUR_FP64, _shared_events(hbl, ps, dt),
                  fill(Int32(ns), hbl.M), dt, newton_tol, itermax)
    for e in ps.dynamic.events
        GradPower.deactivate!(e)
    end
    return collect(0:dt:tf), trajs
end

"""
    integrate_hip_multi!(hbl, ps, tf, fault_buses, rfaults, ton, toff; ...)

`integrate_gpu_shared_multi!` (ext:2640-2708): one fault per scenario
(1-based internal bus numbers, as the reference); `tf`, `ton`, `toff` may be
scalars or per-scenario vectors (uneven horizons).  Returns `(tvecs, trajs)`.
"""
function integrate_hip_multi!(hbl::HipBatchedLayout, ps, tf, fault_buses::AbstractVector{<:Integer},
                              rfaults::AbstractVector{Float64}, ton, toff; dt::Float64 = 1.0 / 120.0,
                              newton_tol::Float64 = 1e-10, itermax::Int = 30, method::Integer = METHOD_SCHUR_FP64)
    M = hbl.M
    @assert length(fault_buses) == M && length(rfaults) == M
    at(x, m) = x isa AbstractVector ? x[m] : x
    ev = [GpEvent(Int32(m - 1), Int32(fault_buses[m] - 1), rfaults[m], Int32(steps(at(ton, m), dt)),
                  Int32(steps(at(toff, m), dt))) for m in 1:M]
    nsteps = Int32[steps(at(tf, m), dt) for m in 1:M]
    trajs = _run!(hbl, ps, Int32(method), ev, nsteps, dt, newton_tol, itermax)
    return [collect(0:Int(nsteps[m])) .* dt for m in 1:M], trajs
end

"""
    integrate_hip_shared!(hbl, ps, tf; dt = 1/120, newton_tol = 1e-10, itermax = 30)

`integrate_gpu_shared!` (ext:2296-2375): one FP64 factorisation of the
reduced network at the initial state shared by every scenario and iteration,
the faults entering each solve as a Woodbury correction; the events of
`ps.dynamic` in every scenario.
"""
function integrate_hip_shared!(hbl::HipBatchedLayout, ps, tf::Float64; dt::Float64 = 1.0 / 120.0,
                               newton_tol::Float64 = 1e-10, itermax::Int = 30)
    ns = steps(tf, dt)
    trajs = _run!(hbl, ps, METHOD_SCHUR_SHARED, _shared_events(hbl, ps, dt), fill(Int32(ns), hbl.M), dt,
                  newton_tol, itermax)
    for e in ps.dynamic.events
        GradPower.deactivate!(e)
    end
    return collect(0:dt:tf), trajs
end

"""
    integrate_hip_shared_multi!(hbl, ps, tf, fault_buses, rfaults, ton, toff; ...)

`integrate_gpu_shared_multi!` (ext:2640-2708): `integrate_hip_multi!` with the
shared factorisation and a rank-two Woodbury correction per scenario.
"""
integrate_hip_shared_multi!(hbl::HipBatchedLayout, ps, tf, fault_buses, rfaults, ton, toff; kw...) =
    integrate_hip_multi!(hbl, ps, tf, fault_buses, rfaults, ton, toff; kw..., method = METHOD_SCHUR_SHARED)

"""
    scenario_status(hbl) -> (status, steps_done)

Outcome of the last integration per scenario: 0 or the error code that stopped
it, and the steps completed before (columns 1..steps_done+1 of its trajectory
are converged states).
"""
function scenario_status(hbl::HipBatchedLayout)
    st = zeros(Int32, hbl.M)
    sd = zeros(Int32, hbl.M)
    check(ccall((:gp_get_scenario_status, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}), hbl.ctx, st, sd))
    return st, sd
end

function symbolic_info(hbl::HipBatchedLayout)
    info = Ref{GpSymbolicInfo}()
    check(ccall((:gp_get_symbolic_info, LIB), Cint, (Ptr{Cvoid}, Ref{GpSymbolicInfo}), hbl.ctx, info))
    return info[]
end

# ---- discrete adjoint ----------------------------------------------------------
"""
    hip_adjoint(hbl, ps, tf, θ, pidx; dt, newton_tol = 1e-12, itermax = 50, segment = 30,
                scale = 1.0, method = METHOD_MONO_FP64) -> (J, dJ/dθ, dJ/dz0)

J = scale * sum_m sum_k dt * sum_g w_g(t_k)^2 over the GENROU speeds (the
reference's `functional`, src/dynamics.jl:330-337, weighted per step as
test/test_adjoint_event_offset.jl:58-69) with pvec[pidx] = θ (1-based indices)
and z0 held fixed (examples/adjoint.jl:45-51); gradients by the discrete
adjoint of the backward-Euler sweep with checkpoints every `segment` steps.
dJ/dz0 is n x M (the λ of the reference's adjoint, per scenario).
"""
function hip_adjoint(hbl::HipBatchedLayout, ps, tf::Float64, θ::AbstractVector{Float64},
                     pidx::AbstractVector{<:Integer}; dt::Float64 = 1.0 / 120.0, newton_tol::Float64 = 1e-12,
                     itermax::Int = 50, segment::Int = 30, scale::Float64 = 1.0,
                     method::Integer = METHOD_MONO_FP64)
    p = copy(hbl.pvec)
    p[pidx] .= θ
    set_state!(hbl, p)
    ev = _shared_events(hbl, ps, dt)
    check(ccall((:gp_set_events, LIB), Cint, (Ptr{Cvoid}, Ptr{GpEvent}, Int32, Ptr{Int32}),
                hbl.ctx, ev, Int32(length(ev)), fill(Int32(steps(tf, dt)), hbl.M)))
    _topology_events!(hbl, ps, dt)
    obj = GpObjective(Int32(0), C_NULL, scale, 0.0)          # rows NULL: the GENROU speeds
    par = Int32.(collect(pidx) .- 1)
    value = zeros(hbl.M)
    grad = zeros(max(1, length(par)))
    dz0 = zeros(hbl.n, hbl.M)                                 # column m: scenario m (row-major [M][n] in C)
    st = Ref{GpAdjointStats}()
    rc = ccall((:gp_adjoint, LIB), Cint,
               (Ptr{Cvoid}, Int32, Float64, Float64, Int32, Int32, Ref{GpObjective}, Int32, Ptr{Int32},
                Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{GpAdjointStats}),
               hbl.ctx, Int32(method), dt, newton_tol, Int32(itermax), Int32(segment), obj, Int32(length(par)),
               par, value, grad, C_NULL, dz0, st)
    set_state!(hbl, hbl.pvec)
    rc == GP_OK || check(rc)
    return sum(value), grad[1:length(par)], dz0
end

hip_objective(hbl, ps, tf, θ, pidx; kw...) = first(hip_adjoint(hbl, ps, tf, θ, pidx; kw...))

# reverse rule: one gp_adjoint call gives the value and the whole gradient
function ChainRulesCore.rrule(::typeof(hip_objective), hbl, ps, tf, θ, pidx; kw...)
    val, g, _ = hip_adjoint(hbl, ps, tf, θ, pidx; kw...)
    pullback(ȳ) = (NoTangent(), NoTangent(), NoTangent(), NoTangent(), ȳ .* g, NoTangent())
    return val, pullback
end

end # module
