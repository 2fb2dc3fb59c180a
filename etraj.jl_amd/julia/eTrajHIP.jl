#=
eTrajHIP.jl -- Julia binding of libetraj_hip.so for eTraj.jl.

Adds a `backend = :hip` path to `perform_traj_simulation` (eTraj.jl
src/TrajectorySimulation.jl:103-161) without touching the Lasers / Targets /
ElectronSamplers types: the Julia objects are flattened into the POD
descriptors of include/etraj.h and the device does the work.

  * fine path   -- `launch_and_collect_hip!(job, batch_id)` replaces the body of
                   `launch_and_collect_2D!/3D!` (TrajectorySimulation.jl:308-500):
                   the Julia sampler still produces the batch, the device integrates
                   and bins it into the thread's `spec_collect` slot.
  * coarse path -- `run_job_hip(; kwargs...)` replaces the whole `@threads` loop
                   and the sum of `write_output!` (:154-157, :521).
  * whole run   -- `perform_traj_simulation_hip(; kwargs...)` takes the reference's
                   keywords, builds its TrajectorySimulationJob (validation, the
                   .jld2 / .h5 file) and writes the same datasets as write_output!.

Usage (in eTraj's module, after `include("TrajectorySimulation.jl")`):
    include("eTrajHIP.jl"); using .eTrajHIP
=#
module eTrajHIP

using ..Lasers
using ..Targets
using ..ElectronSamplers          # gen_electron_batch, batch_num (ElectronSamplers.jl:9)

const libetraj = get(ENV, "ETRAJ_HIP_LIB", joinpath(@__DIR__, "..", "etraj_hip", "libetraj_hip.so"))

# ---- include/etraj.h descriptors (field order and types must match) ----------
struct MonoLaser
    kind::Int32; reserved::Int32
    peak_int::Float64; wave_len::Float64; cyc_num::Float64
    ellip::Float64; azi::Float64; cep::Float64; t_shift::Float64
end
struct LaserDesc
    n_colors::Int32; reserved::Int32; delay21::Float64
    color1::MonoLaser; color2::MonoLaser
end
struct TargetDesc
    kind::Int32; nucl_charge::Int32; l::Int32; m::Int32
    Ip::Float64; asymp_coeff::Float64; quan_ax_theta::Float64; quan_ax_phi::Float64; soft_core::Float64
    a1::Float64; b1::Float64; a2::Float64; b2::Float64; a3::Float64; b3::Float64
end
struct SamplerDesc
    init_method::Int32; dimension::Int32; phase_method::Int32; rate_prefix::Int32
    sampling::Int32; sample_t_num::Int32; ss_kd_num::Int32; ss_kz_num::Int32
    mc_kt_num::Int64
    sample_t_lo::Float64; sample_t_hi::Float64; ss_kd_max::Float64; ss_kz_max::Float64
    mc_kd_max::Float64; mc_kz_max::Float64; kd_sigma::Float64; kz_sigma::Float64
    cutoff_limit::Float64; seed::UInt64
    mol_rot::NTuple{3,Float64}; wfat_na::Int32; wfat_nb::Int32; wfat_coeff::Ptr{Float64}
    wfat_form::Int32; wfat_dims::NTuple{4,Int32}; wfat_reserved::Int32; wfat_mu::NTuple{3,Float64}
    orbit_ip::Float64; asymp_lmax::Int32; asymp_reserved::Int32; asymp_coeff::Ptr{Float64}
end
struct RunDesc
    dimension::Int32; phase_method::Int32; integrator::Int32; precision::Int32
    traj_t_final::Float64; dt::Float64; rtol::Float64; atol::Float64; restep_radius::Float64
    final_p_max::NTuple{3,Float64}; final_p_num::NTuple{3,Int32}; quirks::Int32; maxiters::Int64
    pmd_mode::Int32; pmd_binning::Int32; hit_counts::Ptr{UInt64}
end
mutable struct TuningDesc   # etraj_tuning: the launch choices of a ctx
    bin_tiles::Int32; bin_tile_shift::Int32; bin_chunk::Int32; overlap_fp64::Int32
    rk4_block::Int32; tsit5_refill::Int32; tsit5_claim::Int32; graph::Int32
    TuningDesc() = new(-1, 0, 16384, 1, 256, 4, 256, 0)
end
mutable struct Stats
    ion_prob::Float64; ion_prob_uncollected::Float64
    num_effective_traj::Int64; n_failed::Int64; n_bound::Int64; n_restep::Int64; n_steps::Int64; n_steps_fp64::Int64
    ms_sample::Float64; ms_integrate::Float64; ms_integrate_main::Float64; ms_total::Float64; n_slabs::Int64
    n_attempts::Int64
    Stats() = new(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
end

const PHASE = Dict(:CTMC => Int32(0), :QTMC => Int32(1), :SCTS => Int32(2))
const BINNING = Dict(:auto => Int32(0), :atomic => Int32(1), :tiles => Int32(2))

mono(l::Cos2Laser) = MonoLaser(1, 0, PeakInt(l), WaveLen(l), CycNum(l), Ellipticity(l), Azimuth(l), CEP(l), TimeShift(l))
mono(l::Cos4Laser) = MonoLaser(2, 0, PeakInt(l), WaveLen(l), CycNum(l), Ellipticity(l), Azimuth(l), CEP(l), TimeShift(l))
mono(l::GaussianLaser) = MonoLaser(3, 0, PeakInt(l), WaveLen(l), SpreadCycNum(l), Ellipticity(l), Azimuth(l), CEP(l), TimeShift(l))
const NOLASER = MonoLaser(0, 0, 0, 0, 0, 0, 0, 0, 0)
laser_desc(l::MonochromaticLaser) = LaserDesc(1, 0, 0.0, mono(l), NOLASER)
laser_desc(l::BichromaticLaser) = LaserDesc(2, 0, Delay21(l), mono(Laser1(l)), mono(Laser2(l)))

target_desc(t::HydrogenLikeAtom) = TargetDesc(1, AsympNuclCharge(t), AngularQuantumNumber(t), MagneticQuantumNumber(t),
    IonPotential(t), AsympCoeff(t), QuantizationAxisOrientaion(t)..., SoftCore(t), 0, 0, 0, 0, 0, 0)
target_desc(t::SAEAtom) = TargetDesc(2, AsympNuclCharge(t), AngularQuantumNumber(t), MagneticQuantumNumber(t),
    IonPotential(t), AsympCoeff(t), QuantizationAxisOrientaion(t)..., SoftCore(t), t.a1, t.b1, t.a2, t.b2, t.a3, t.b3)
# GenericMolecule: the single-centre soft-core potential -(Q+1)/sqrt(r^2+1) with the HOMO Ip
# (GenericMolecule.jl:450-457), i.e. a hydrogen-like descriptor with soft core 1
target_desc(t::GenericMolecule) = TargetDesc(1, AsympNuclCharge(t), 0, 0, IonPotential(t), 1.0, 0.0, 0.0, 1.0,
    0, 0, 0, 0, 0, 0)

# pmd_mode :fixed = ETRAJ_PMD_FIXED (order-independent 128-bit fixed-point sums); hit_counts: C_NULL or
# the pointer of a Vector{UInt64} of prod(final_p_num) cells (kept alive by the caller) that receives
# the number of electrons binned into each cell
# integrator :rk4 (fixed step dt) or :tsit5 (adaptive, reltol `rtol`, abstol 1e-6 as OrdinaryDiffEq's default)
# pmd_binning :auto, :atomic or :tiles (ETRAJ_BIN_*: LDS-privatised PMD tiles)
function run_desc(dimension, phase, traj_t_final, final_p_max, final_p_num; dt=0.1, precision=:mixed,
                  restep_radius=2.0, reproduce_quirks=true, pmd_mode=:float, hit_counts=Ptr{UInt64}(C_NULL),
                  integrator=:rk4, rtol=1e-6, pmd_binning=:auto)
    pmax = ntuple(i -> i <= dimension ? Float64(final_p_max[i]) : 1.0, 3)
    pnum = ntuple(i -> i <= dimension ? Int32(final_p_num[i]) : Int32(1), 3)
    integ = integrator == :tsit5 ? Int32(1) : integrator == :rk4 ? Int32(0) :
            error("[eTrajHIP] unknown integrator $integrator")
    RunDesc(dimension, PHASE[phase], integ, precision == :mixed ? 1 : 0, traj_t_final, dt, rtol, 1e-6, restep_radius,
            pmax, pnum, reproduce_quirks ? 3 : 0, 1_000_000, pmd_mode == :fixed ? 1 : 0,
            Int32(BINNING[pmd_binning]), hit_counts)
end

# ---- context ------------------------------------------------------------------
mutable struct Ctx
    h::Ptr{Cvoid}
end
function Ctx(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:etraj_ctx_create, libetraj), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r)
    rc == 0 || error("[eTrajHIP] etraj_ctx_create failed ($rc): no usable HIP device")
    c = Ctx(r[])
    finalizer(x -> ccall((:etraj_ctx_destroy, libetraj), Cvoid, (Ptr{Cvoid},), x.h), c)
    return c
end
function check(c::Ctx, rc::Cint)
    rc == 0 && return
    msg = unsafe_string(ccall((:etraj_last_error, libetraj), Cstring, (Ptr{Cvoid},), c.h))
    rc == -1 ? throw(AssertionError(msg)) : error("[eTrajHIP] $msg (code $rc)")
end

# One ctx per Julia thread: a ctx is single-threaded (include/etraj.h).
const CTXS = Dict{Int,Ctx}()
const CTX_LOCK = ReentrantLock()
thread_ctx() = lock(CTX_LOCK) do
    get!(() -> Ctx(0), CTXS, Threads.threadid())
end

# ---- fine path: drop-in for the per-batch body of launch_and_collect! ----------
function launch_and_collect_hip!(job, batch_id::Integer; dt=0.1, precision=:mixed, restep_radius=2.0,
                                 integrator=:rk4)
    (job.batch_completion[batch_id] == 1) && return
    init = gen_electron_batch(job.sp, batch_id)                         # the reference sampler, unchanged
    isnothing(init) && (job.batch_completion[batch_id] = 1; return)
    tid = Threads.threadid()
    rd = Ref(run_desc(job.dimension, job.traj_phase_method, job.traj_t_final, job.final_p_max, job.final_p_num;
                      dt, precision, restep_radius, integrator, rtol=job.traj_rtol))
    ld = Ref(laser_desc(job.laser)); td = Ref(target_desc(job.target))
    ip = Ref(0.0); iu = Ref(0.0); nf = Ref(Int64(0))
    spec = job.spec_collect[tid]                                        # Float64 or ComplexF64 array, px fastest
    c = thread_ctx()
    GC.@preserve init spec begin
        rc = ccall((:etraj_integrate_and_bin, libetraj), Cint,
                   (Ptr{Cvoid}, Ref{LaserDesc}, Ref{TargetDesc}, Ref{RunDesc}, Ptr{Float64}, Int64, Int32,
                    Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Int64}),
                   c.h, ld, td, rd, init, size(init, 2), size(init, 1), spec, ip, iu, nf)
        check(c, rc)
    end
    job.classical_prob[tid] += ip[]
    job.classical_prob_uncollected[tid] += iu[]
    nf[] > 0 && @warn "[launch_and_collect_hip!] $(nf[]) electrons failed in batch $batch_id."
    job.batch_completion[batch_id] = 1
    job.eff_trajs[batch_id] += size(init, 2)
    return
end

# ---- coarse path: the whole loop on the device ----------------------------------
prefix_bits(p::Symbol) = p == :Exp ? 0 : p == :Full ? 6 : p == :Pre ? 1 : p == :PreCC ? 2 : p == :Jac ? 4 :
                         error("[ADKSampler] Undefined tunneling rate prefix [$p].")
prefix_bits(p) = isempty(p) ? 0 : reduce(|, prefix_bits.(collect(p)))

# init_cond_method -> ETRAJ_INIT_* (ElectronSamplers.jl:11-24 aliases)
init_method(m::Symbol) = m in (:ADK, :MOADK) ? Int32(1) : m in (:SPA, :MOSPA, :SFA, :MOSFA) ? Int32(2) :
                         m in (:SPANE, :MOSPANE, :SFAAE, :MOSFAAE) ? Int32(3) : m == :WFAT ? Int32(4) :
                         error("[ElectronSamplers] Undefined initial condition method [$m].")

# The sampler descriptor and the arrays it points into (keep them under GC.@preserve).
# WFAT: every channel of the WFAT integrals MolWFATData(mol, ridx) (GenericMolecule.jl:228-233),
# passed as Julia's own Array{ComplexF64,4} memory (ETRAJ_WFAT_INTDATA = 1).
# ADK / SPA / SPANE of a molecule (MO-ADK, MO-SFA): MolAsympCoeff(mol, ridx), Julia's own
# Matrix{ComplexF64} memory. Atoms with l > 0 need nothing here (the target descriptor's l, m).
# sample_monte_carlo = true: sorted random t0 and uniform kd [kz] as ADKSampler.jl:140-147, drawn
# from Philox4x64-10 keyed by `seed` (the reference's unseeded rand() cannot be reproduced).
function sampler_desc(; init_cond_method=:ADK, dimension, sample_t_intv, sample_t_num, ss_kd_max=0.0, ss_kd_num=0,
                      ss_kz_max=0.0, ss_kz_num=0, traj_phase_method=:CTMC, sample_cutoff_limit=1e-16,
                      rate_prefix=:Full, target=nothing, mol_orbit_ridx=0, sample_monte_carlo=false,
                      mc_kt_num=0, mc_kd_max=0.0, mc_kz_max=0.0, seed=0, kwargs...)
    rot = (0.0, 0.0, 0.0); form = Int32(0); dims = (Int32(0), Int32(0), Int32(0), Int32(0))
    mu = (0.0, 0.0, 0.0); ip = 0.0; coeff = C_NULL; keep = nothing; na = Int32(0); nb = Int32(0)
    alm = Int32(0); acoeff = C_NULL
    if init_cond_method != :WFAT && target isa GenericMolecule
        C = MolAsympCoeff(target, mol_orbit_ridx)               # GenericMolecule.jl:304-309
        keep = C
        acoeff = Ptr{Float64}(pointer(C)); alm = Int32(size(C, 1) - 1)
        ip = IonPotential(target, mol_orbit_ridx); rot = Float64.(MolRotation(target))
    elseif init_cond_method == :WFAT
        μ, intdata = MolWFATData(target, mol_orbit_ridx)
        keep = intdata
        coeff = Ptr{Float64}(pointer(intdata))
        form = Int32(1); dims = Int32.(size(intdata)); mu = Tuple(Float64.(μ))
        ip = IonPotential(target, mol_orbit_ridx); rot = Float64.(MolRotation(target))
        na, nb = dims[3], dims[4]
    end
    sd = SamplerDesc(init_method(init_cond_method), dimension, PHASE[traj_phase_method], prefix_bits(rate_prefix),
                     sample_monte_carlo ? 1 : 0, sample_t_num, ss_kd_num, dimension == 3 ? ss_kz_num : 1,
                     sample_monte_carlo ? mc_kt_num : 0, sample_t_intv[1], sample_t_intv[2],
                     ss_kd_max, ss_kz_max, mc_kd_max, dimension == 3 ? mc_kz_max : 0.0, 0, 0, sample_cutoff_limit,
                     UInt64(seed), rot, na, nb, coeff, form, dims, 0, mu, ip, alm, 0, acoeff)
    return sd, keep
end

function run_job_hip(; laser, target, dimension, sample_t_intv, sample_t_num, traj_t_final, final_p_max, final_p_num,
                     ss_kd_max=0.0, ss_kd_num=0, ss_kz_max=0.0, ss_kz_num=0, traj_phase_method=:CTMC,
                     sample_cutoff_limit=1e-16, rate_prefix=:Full, dt=0.1, precision=:mixed, restep_radius=2.0,
                     init_cond_method=:ADK, mol_orbit_ridx=0, device=0, integrator=:rk4, traj_rtol=1e-6,
                     pmd_mode=:float, kwargs...)
    sdv, keep = sampler_desc(; init_cond_method, dimension, sample_t_intv, sample_t_num, ss_kd_max, ss_kd_num,
                             ss_kz_max, ss_kz_num, traj_phase_method, sample_cutoff_limit, rate_prefix, target,
                             mol_orbit_ridx, kwargs...)
    sd = Ref(sdv)
    rd = Ref(run_desc(dimension, traj_phase_method, traj_t_final, final_p_max, final_p_num; dt, precision, restep_radius,
                      integrator, rtol=traj_rtol, pmd_mode))
    spec = zeros(Float64, final_p_num...)
    st = Stats()
    c = Ctx(device)
    GC.@preserve spec st keep begin
        rc = ccall((:etraj_run_job, libetraj), Cint,
                   (Ptr{Cvoid}, Ref{LaserDesc}, Ref{TargetDesc}, Ref{SamplerDesc}, Ref{RunDesc}, Ptr{Float64}, Ref{Stats}),
                   c.h, Ref(laser_desc(laser)), Ref(target_desc(target)), sd, rd, spec, st)
        check(c, rc)
    end
    return (momentum_spec = spec, ion_prob = st.ion_prob, ion_prob_uncollected = st.ion_prob_uncollected,
            num_effective_traj = st.num_effective_traj)
end

# ---- launch tuning: etraj_autotune times the ctx's launch choices on a job ----------
# (the keywords of run_job_hip); the fastest set stays in the ctx `c` for its later calls.
function autotune_hip(c::Ctx; laser, target, dimension, sample_t_intv, sample_t_num, traj_t_final, final_p_max,
                      final_p_num, traj_phase_method=:CTMC, dt=0.1, precision=:mixed, restep_radius=2.0,
                      integrator=:rk4, traj_rtol=1e-6, pmd_mode=:float, pmd_binning=:auto, kwargs...)
    sdv, keep = sampler_desc(; dimension, sample_t_intv, sample_t_num, traj_phase_method, target, kwargs...)
    sd = Ref(sdv)
    rd = Ref(run_desc(dimension, traj_phase_method, traj_t_final, final_p_max, final_p_num; dt, precision,
                      restep_radius, integrator, rtol=traj_rtol, pmd_mode, pmd_binning))
    best = TuningDesc()
    best_ms, start_ms = Ref{Float64}(0), Ref{Float64}(0)
    GC.@preserve keep best begin
        rc = ccall((:etraj_autotune, libetraj), Cint,
                   (Ptr{Cvoid}, Ref{LaserDesc}, Ref{TargetDesc}, Ref{SamplerDesc}, Ref{RunDesc}, Ref{TuningDesc},
                    Ref{Float64}, Ref{Float64}),
                   c.h, Ref(laser_desc(laser)), Ref(target_desc(target)), sd, rd, best, best_ms, start_ms)
        check(c, rc)
    end
    return (best = best, best_ms = best_ms[], start_ms = start_ms[])
end

# ---- parameter scans: the reference's loops of perform_traj_simulation calls --------
# (examples/test_Bichromatic_CCP.jl:4-27) as one etraj_run_scan; `lasers` and the
# sampler keyword sets differ per point, target and run settings are shared.
function run_scan_hip(lasers::AbstractVector; target, dimension, sample_t_intv, sample_t_num, traj_t_final,
                      final_p_max, final_p_num, traj_phase_method=:CTMC, dt=0.1, precision=:mixed,
                      restep_radius=2.0, device=0, integrator=:rk4, traj_rtol=1e-6, pmd_mode=:float, kwargs...)
    n = length(lasers)
    lds = [laser_desc(l) for l in lasers]
    sdk = [sampler_desc(; dimension, sample_t_intv, sample_t_num, traj_phase_method, target, kwargs...) for _ in 1:n]
    sds = first.(sdk); keeps = last.(sdk)
    rd = Ref(run_desc(dimension, traj_phase_method, traj_t_final, final_p_max, final_p_num; dt, precision, restep_radius,
                      integrator, rtol=traj_rtol, pmd_mode))
    npix = prod(final_p_num)
    specs = zeros(Float64, npix, n)
    stv = Vector{NTuple{13,Float64}}(undef, n)   # etraj_stats is 13 eight-byte fields
    c = Ctx(device)
    GC.@preserve lds sds keeps specs stv begin
        rc = ccall((:etraj_run_scan, libetraj), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{LaserDesc}, Ref{TargetDesc}, Ptr{SamplerDesc}, Ref{RunDesc}, Ptr{Float64},
                    Ptr{Cvoid}),
                   c.h, n, lds, Ref(target_desc(target)), sds, rd, specs, stv)
        check(c, rc)
    end
    return [reshape(specs[:, j], final_p_num...) for j in 1:n]
end

# ---- whole-run drop-in for perform_traj_simulation (TrajectorySimulation.jl:103-161) ----
# The reference's TrajectorySimulationJob validates the keywords, opens `output_path!` as .jld2
# or .h5 and records `params`, so the file is written by JLD2.jl / HDF5.jl themselves.
#   coarse = true   etraj_run_job samples, integrates and bins on the device; the result is
#                   written by write_output_hip! (the dataset set of write_output!, :502-546,
#                   with the PMD that already holds abs2 for QTMC / SCTS).
#   coarse = false  the reference sampler makes every batch and launch_and_collect_hip!
#                   integrates it; the reference's own write_output! writes the file.
# Keywords and defaults are the reference's (:105-137) plus the device settings: dt, precision
# (:mixed / :fp64), restep_radius, seed, device, integrator (:rk4, or :tsit5 at traj_rtol, the
# reference's integrator) and pmd_mode (:float / :fixed).
function perform_traj_simulation_hip(; init_cond_method::Symbol, laser::Laser, target::Target, dimension::Integer,
        sample_t_intv, sample_t_num::Integer, traj_t_final, final_p_max, final_p_num,
        ss_kd_max::Real=0.0, ss_kd_num::Integer=0, ss_kz_max::Real=0.0, ss_kz_num::Integer=0,
        mc_kt_num::Integer=0, mc_kd_max::Real=0.0, mc_kz_max::Real=0.0,
        traj_phase_method::Symbol=:CTMC, traj_rtol::Real=1e-6, output_fmt::Symbol=:jld2,
        output_compress::Bool=true, output_path::String=parentmodule(@__MODULE__).default_filename(),
        sample_monte_carlo::Bool=false, sample_cutoff_limit::Real=1e-16, rate_prefix=:Full, mol_orbit_ridx=0,
        show_progress::Bool=false, coarse::Bool=true, dt=0.1, precision=:mixed, restep_radius=2.0, seed=0,
        device=0, integrator=:rk4, pmd_mode=:float)
    ref = parentmodule(@__MODULE__)
    job = ref.TrajectorySimulationJob(; init_cond_method, laser, target, dimension, sample_t_intv, sample_t_num,
        traj_t_final, final_p_max, final_p_num, ss_kd_max, ss_kd_num, ss_kz_max, ss_kz_num, mc_kt_num, mc_kd_max,
        mc_kz_max, traj_phase_method, traj_rtol, output_fmt, output_compress, output_path, sample_monte_carlo,
        sample_cutoff_limit, rate_prefix, mol_orbit_ridx)
    if coarse
        res = run_job_hip(; laser, target, dimension, sample_t_intv, sample_t_num, traj_t_final, final_p_max,
                          final_p_num, ss_kd_max, ss_kd_num, ss_kz_max, ss_kz_num, traj_phase_method,
                          sample_cutoff_limit, rate_prefix, dt, precision, restep_radius, init_cond_method,
                          mol_orbit_ridx, device, integrator, traj_rtol, pmd_mode, sample_monte_carlo, mc_kt_num,
                          mc_kd_max, mc_kz_max, seed)
        write_output_hip!(job, res)
    else
        Threads.@threads for i in 1:batch_num(job.sp)
            launch_and_collect_hip!(job, i; dt, precision, restep_radius, integrator)
        end
        ref.write_output!(job)
    end
    @info "Task finished, data saved at `$(job.output_path)`."
    return job.output_path
end

# write_output! (:502-546) for a PMD reduced on the device. The keys, the params / params_text
# encoding and the compression settings are the reference's, so readers such as
# examples/plot_1cyc_LP.jl:9-15 see the same file.
function write_output_hip!(job, res)
    ref = parentmodule(@__MODULE__)
    file = job.file
    params = copy(job.params)
    params[:laser] = Lasers.Serialize(params[:laser])
    params[:target] = Targets.Serialize(params[:target])
    file["params_text"] = ref.YAML.write(params)
    job.output_fmt == :jld2 && (file["params"] = job.params)
    file["px"] = job.px
    file["py"] = job.py
    job.dimension == 3 && (file["pz"] = job.pz)
    spec = res.momentum_spec                          # Float64, abs2 taken for QTMC / SCTS
    if job.output_fmt == :jld2
        write(file, "momentum_spec", spec; compress=job.output_compress)   # JLD2's method of Base.write
    elseif job.output_compress
        file["momentum_spec", shuffle=true, deflate=6] = spec
    else
        file["momentum_spec"] = spec
    end
    file["ion_prob"] = res.ion_prob
    file["ion_prob_uncollected"] = res.ion_prob_uncollected
    file["num_effective_traj"] = res.num_effective_traj
    close(file)
    mv(job.output_path * "!", job.output_path)
    return job.output_path
end

export launch_and_collect_hip!, run_job_hip, run_scan_hip, perform_traj_simulation_hip, autotune_hip

end # module
