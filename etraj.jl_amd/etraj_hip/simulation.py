"""perform_traj_simulation with the HIP engine (TrajectorySimulation.jl:103-161).

Keyword interface of the reference, plus engine keywords:
  traj_integrator  "tsit5" (default): the reference's adaptive Tsit5 at traj_rtol
                   (TrajectorySimulation.jl:342), in FP64 unless precision says
                   otherwise; "rk4": fixed-step RK4 at traj_dt, mixed precision by
                   default (the engine's throughput form; its PMD differs from the
                   reference algorithm's by its truncation error, and traj_rtol
                   does not apply to it -- a warning says so if it is given)
  traj_dt          RK4 nominal step (a.u., default 0.1)
  precision        "mixed" (FP32 EOM, FP64 phase/Kepler/re-step) or "fp64"
  restep_radius    FP64 re-step radius of the mixed mode (a.u., default 2)
  granularity      "job": the whole batch loop on the device (etraj_run_job);
                   "batch": per batch gen_electron_batch + integrate_and_bin,
                   the drop-in for launch_and_collect! (:300-401)
  seed, sample_stratified, kd_sigma, kz_sigma   seeded sampling modes
  reproduce_quirks reproduce the reference's 3D SCTS / SAE 2D SCTS formulas
  pmd_accumulator  "float": FP64 atomic sums (the reference's Float64 sums, rounded
                   in arrival order); "fixed": 128-bit fixed-point sums (2^-100
                   resolution), bit-identical for any order of the electrons
  pmd_binning      "auto" (the context's choice), "atomic" (one atomic addition per
                   electron from the integrator) or "tiles" (electrons sorted by PMD
                   tile, each tile accumulated in LDS and added to the PMD once)
  hit_counts       also return result["hit_counts"]: the number of electrons
                   binned into each PMD cell (the integer-count PMD)
Returns the fields write_output! stores (:516-543) as a dict; when
output_path is given they are also saved: output_fmt "jld2" (the reference's
default) writes the JLD2 file and "h5" / "hdf5" the HDF5 file of write_output!
(etraj_hip.h5out; the JLD2 file lacks only `params`, the parameters as Julia
objects, which params_text holds as YAML), "npz" a numpy archive of the same
fields.
"""
from __future__ import annotations

import os
import time
import warnings

import numpy as np

from . import _abi
from .lasers import LaserFx, LaserFy, Serialize as _ser_laser
from .samplers import init_sampler, batch_num, gen_electron_batch, _sym
from .targets import Serialize as _ser_target


def run_desc(*, dimension, traj_phase_method, traj_t_final, final_p_max, final_p_num, traj_integrator="rk4",
             traj_dt=0.1, traj_rtol=1e-6, precision=None, restep_radius=2.0, reproduce_quirks=True,
             pmd_accumulator="float", pmd_binning="auto", **_):
    d = _abi.RunDesc()
    d.dimension = int(dimension)
    d.phase_method = _abi.PHASE_METHODS[_sym(traj_phase_method)]
    integ = _sym(traj_integrator).lower()
    if integ not in ("rk4", "tsit5"):
        raise ValueError(f"unknown traj_integrator {traj_integrator}")
    d.integrator = _abi.INTEG_RK4 if integ == "rk4" else _abi.INTEG_TSIT5
    if precision is None:   # mixed for the fixed-step RK4, FP64 for Tsit5 (its only form)
        precision = "mixed" if integ == "rk4" else "fp64"
    prec = _sym(precision).lower()
    if prec not in ("mixed", "fp64"):
        raise ValueError(f"unknown precision {precision}")
    d.precision = _abi.PREC_MIXED if prec == "mixed" else _abi.PREC_FP64
    d.traj_t_final = float(traj_t_final)
    d.dt = float(traj_dt)
    d.rtol = float(traj_rtol)
    d.atol = 1e-6
    d.restep_radius = float(restep_radius)
    for k in range(dimension):
        d.final_p_max[k] = float(final_p_max[k])
        d.final_p_num[k] = int(final_p_num[k])
    for k in range(dimension, 3):
        d.final_p_max[k] = 1.0
        d.final_p_num[k] = 1
    d.quirks = _abi.QUIRKS_REFERENCE if reproduce_quirks else 0
    d.maxiters = 1000000
    acc = _sym(pmd_accumulator).lower()
    if acc not in ("float", "fixed"):
        raise ValueError(f"unknown pmd_accumulator {pmd_accumulator}")
    d.pmd_mode = _abi.PMD_FIXED if acc == "fixed" else _abi.PMD_FLOAT
    binning = {"auto": _abi.BIN_AUTO, "atomic": _abi.BIN_ATOMIC, "tiles": _abi.BIN_TILES}
    if _sym(pmd_binning).lower() not in binning:
        raise ValueError(f"unknown pmd_binning {pmd_binning}")
    d.pmd_binning = binning[_sym(pmd_binning).lower()]
    return d


def _check_rtol(traj_integrator, traj_rtol):
    """traj_rtol of the reference (default 1e-6); a warning when it is given to the fixed-step RK4,
    which does not use it."""
    if traj_rtol is None:
        return 1e-6
    if _sym(traj_integrator).lower() == "rk4":
        warnings.warn("[perform_traj_simulation] traj_rtol applies to the adaptive Tsit5 integrator (the reference's "
                      "algorithm); the fixed-step RK4 integrates at traj_dt and ignores it.")
    return traj_rtol


def _axis(pmax, num):
    return np.linspace(-pmax, pmax, num) if num > 1 else np.array([0.0])


def default_filename():
    return time.strftime("TrajectorySimulation-%Y%m%d.%H%M%S")


def perform_traj_simulation(*, init_cond_method, laser, target, dimension, sample_t_intv, sample_t_num,
                            traj_t_final, final_p_max, final_p_num, ss_kd_max=0.0, ss_kd_num=0, ss_kz_max=0.0,
                            ss_kz_num=0, mc_kt_num=0, mc_kd_max=0.0, mc_kz_max=0.0, traj_phase_method="CTMC",
                            traj_rtol=None, output_fmt="jld2", output_compress=True, output_path=None,
                            sample_monte_carlo=False, sample_cutoff_limit=1e-16, rate_prefix="Full",
                            mol_orbit_ridx=0, show_progress=False, traj_integrator="tsit5", traj_dt=0.1,
                            precision=None, restep_radius=2.0, granularity="job", seed=0,
                            sample_stratified=False, kd_sigma=0.0, kz_sigma=0.0, reproduce_quirks=True,
                            pmd_accumulator="float", pmd_binning="auto", hit_counts=False, backend="hip"):
    if _sym(backend) != "hip":
        raise ValueError("this package is the HIP backend; the CPU method is the reference itself")
    traj_rtol = _check_rtol(traj_integrator, traj_rtol)
    # TrajectorySimulationJob checks (:175-186)
    assert sample_t_intv[0] < sample_t_intv[1], "[TrajectorySimulationJob] `sample_t_intv` must be a valid interval."
    assert traj_t_final > sample_t_intv[1], \
        "[TrajectorySimulationJob] `traj_t_final` must be greater than the last time point of `sample_t_intv`."
    assert dimension in (2, 3), "[TrajectorySimulationJob] `dimension` must be either 2 or 3."
    assert len(final_p_max) == len(final_p_num) == dimension, \
        "[TrajectorySimulationJob] `length(final_p_max)` and `length(final_p_num)` should match `dimension`."
    Fx, Fy = LaserFx(laser), LaserFy(laser)
    for tt, what in ((sample_t_intv[0], "beginning of the sampling time interval"),
                     (sample_t_intv[1], "end of the sampling time interval"), (traj_t_final, "end of the trajectory simulation")):
        if Fx(tt) ** 2 + Fy(tt) ** 2 > 0.005 ** 2:
            warnings.warn(f"[TrajectorySimulationJob] The laser field is not vanishingly small at the {what}.")
    sp = init_sampler(init_cond_method=init_cond_method, laser=laser, target=target, dimension=dimension,
                      sample_t_intv=sample_t_intv, sample_t_num=sample_t_num, sample_cutoff_limit=sample_cutoff_limit,
                      sample_monte_carlo=sample_monte_carlo, traj_phase_method=traj_phase_method,
                      rate_prefix=rate_prefix, ss_kd_max=ss_kd_max, ss_kd_num=ss_kd_num, ss_kz_max=ss_kz_max,
                      ss_kz_num=ss_kz_num, mc_kt_num=mc_kt_num, mc_kd_max=mc_kd_max, mc_kz_max=mc_kz_max,
                      mol_orbit_ridx=mol_orbit_ridx, seed=seed, sample_stratified=sample_stratified,
                      kd_sigma=kd_sigma, kz_sigma=kz_sigma)
    rd = run_desc(dimension=dimension, traj_phase_method=traj_phase_method, traj_t_final=traj_t_final,
                  final_p_max=final_p_max, final_p_num=final_p_num, traj_integrator=traj_integrator,
                  traj_dt=traj_dt, traj_rtol=traj_rtol, precision=precision, restep_radius=restep_radius,
                  reproduce_quirks=reproduce_quirks, pmd_accumulator=pmd_accumulator, pmd_binning=pmd_binning)
    ctx = _abi.default_context()
    npix = int(np.prod(final_p_num))
    hits = np.zeros(npix, dtype=np.uint64) if hit_counts else None
    if hits is not None:
        rd.hit_counts = hits.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64))
    coherent = rd.phase_method != _abi.PHASE_CTMC
    stats = {}
    if granularity == "job":
        spec = np.empty(npix)
        st = _abi.Stats()
        ctx.check(ctx.lib.etraj_run_job(ctx.handle, laser._desc(), target._desc(), sp.desc, rd, _abi.dptr(spec),
                                        _abi.C.byref(st)))
        stats = st.as_dict()
        ion_prob, ion_unc, n_eff = st.ion_prob, st.ion_prob_uncollected, st.num_effective_traj
    elif granularity == "batch":
        acc = np.zeros(2 * npix if coherent else npix)
        ip, iu = _abi.C.c_double(0.0), _abi.C.c_double(0.0)
        nf = _abi.C.c_int64(0)
        n_eff = 0
        for b in range(1, batch_num(sp) + 1):
            init = gen_electron_batch(sp, b)
            if init is None:
                continue
            cols = np.ascontiguousarray(init.T)
            ctx.check(ctx.lib.etraj_integrate_and_bin(ctx.handle, laser._desc(), target._desc(), rd, _abi.dptr(cols),
                                                      cols.shape[0], cols.shape[1], acc.ctypes.data_as(_abi.C.c_void_p),
                                                      _abi.C.byref(ip), _abi.C.byref(iu), _abi.C.byref(nf)))
            n_eff += cols.shape[0]
        spec = (acc[0::2] ** 2 + acc[1::2] ** 2) if coherent else acc
        ion_prob, ion_unc = ip.value, iu.value
    else:
        raise ValueError("granularity must be 'job' or 'batch'")
    result = _result(spec, ion_prob, ion_unc, n_eff, stats, init_cond_method=init_cond_method, laser=laser,
                     target=target, dimension=dimension, sample_t_intv=sample_t_intv, sample_t_num=sample_t_num,
                     traj_phase_method=traj_phase_method, traj_t_final=traj_t_final, final_p_max=final_p_max,
                     final_p_num=final_p_num, traj_integrator=traj_integrator, traj_dt=traj_dt, traj_rtol=traj_rtol,
                     precision=precision)
    if hits is not None:
        result["hit_counts"] = _julia_shape(hits, final_p_num)
    _save(result, output_path, output_fmt, output_compress)
    return result


def _julia_shape(a, final_p_num):
    shape = tuple(int(n) for n in final_p_num)
    return a.reshape(shape[::-1]).T


def _result(spec, ion_prob, ion_unc, n_eff, stats, **kw):
    """The fields write_output! stores (:516-543); PMD layout px fastest (Julia column-major)."""
    final_p_max, final_p_num = kw["final_p_max"], kw["final_p_num"]
    result = {
        "px": _axis(final_p_max[0], final_p_num[0]),
        "py": _axis(final_p_max[1], final_p_num[1]),
        "momentum_spec": _julia_shape(spec, final_p_num),
        "ion_prob": ion_prob,
        "ion_prob_uncollected": ion_unc,
        "num_effective_traj": n_eff,
        "params": {k: kw.get(k) for k in ("init_cond_method", "sample_t_intv", "sample_t_num", "traj_phase_method",
                                          "traj_t_final", "final_p_max", "final_p_num", "traj_integrator", "traj_dt",
                                          "traj_rtol", "precision")},
        "stats": stats,
    }
    result["params"]["laser"] = _ser_laser(kw["laser"])
    result["params"]["target"] = _ser_target(kw["target"])
    if kw["dimension"] == 3:
        result["pz"] = _axis(final_p_max[2], final_p_num[2])
    return result


def _params_text(params) -> str:
    """YAML text of the run parameters (write_output!'s params_text, :507-511)."""
    import yaml

    def plain(v):
        if isinstance(v, dict):
            return {str(k): plain(x) for k, x in v.items()}
        if isinstance(v, (list, tuple)):
            return [plain(x) for x in v]
        if isinstance(v, np.ndarray):
            return plain(v.tolist())
        if isinstance(v, (np.integer,)):
            return int(v)
        if isinstance(v, (np.floating,)):
            return float(v)
        if isinstance(v, (bool, int, float, str)) or v is None:
            return v
        return str(v)
    return yaml.safe_dump(plain(params), sort_keys=False, allow_unicode=True)


def _save(result, output_path, output_fmt="jld2", output_compress=True):
    """The file handling of TrajectorySimulationJob / write_output! (:195-215,
    :502-546): a suffix per format, a default file name instead of an existing
    file or temp file, the data written to `path!` and renamed at the end."""
    if output_path is None:
        return
    fmt = _sym(output_fmt).lower()
    if fmt in ("h5", "hdf5"):
        ext = ".h5"
    elif fmt == "npz":
        ext = ".npz"
    elif fmt == "jld2":
        ext = ".jld2"
    else:
        raise ValueError(f"[TrajectorySimulationJob] Unsupported output format [{output_fmt}].")
    path = output_path if output_path.endswith(ext) else output_path + ext
    if os.path.exists(path):
        warnings.warn(f"[TrajectorySimulationJob] File `{path}` already exists, will save at default filename.")
        path = default_filename() + ext
    if os.path.exists(path + "!"):
        warnings.warn(f"[TrajectorySimulationJob] The temp file `{path}!` exists, will save at default filename.")
        path = default_filename() + ext
    if ext in (".h5", ".jld2"):
        from .h5out import write_h5
        write_h5(path, result, _params_text(result["params"]), compress=bool(output_compress), jld2=ext == ".jld2")
    else:
        tmp = path + "!"
        with open(tmp, "wb") as fh:
            np.savez_compressed(fh, **{k: v for k, v in result.items() if isinstance(v, (np.ndarray, float, int))})
        os.replace(tmp, path)
    result["output_path"] = path


def autotune(**kwargs) -> dict:
    """The engine's autotuner (etraj_autotune) on the job that perform_traj_simulation(**kwargs) would run:
    each candidate launch choice (workgroup size, binning, tile size, FP64-pass overlap for RK4; refill
    threshold, queue claim, binning for Tsit5) is timed on the whole job, and the fastest set stays in the
    default context for the calls that follow. Returns {"best": choices, "best_ms": ms per job, "start_ms": ms
    per job with the choices it started from}."""
    kw = dict(_SCAN_DEFAULTS)
    kw.update(kwargs)
    kw["traj_rtol"] = _check_rtol(kw["traj_integrator"], kw["traj_rtol"])
    sp = init_sampler(**{k: v for k, v in kw.items() if k not in ("output_path", "output_fmt", "output_compress",
                                                                   "hit_counts")})
    rd = run_desc(**kw)
    ctx = _abi.default_context()
    best = _abi.Tuning()
    best_ms, start_ms = _abi.C.c_double(0), _abi.C.c_double(0)
    ctx.check(ctx.lib.etraj_autotune(ctx.handle, kw["laser"]._desc(), kw["target"]._desc(), sp.desc, rd,
                                     _abi.C.byref(best), _abi.C.byref(best_ms), _abi.C.byref(start_ms)))
    return {"best": best.as_dict(), "best_ms": best_ms.value, "start_ms": start_ms.value}


_SCAN_DEFAULTS = dict(ss_kd_max=0.0, ss_kd_num=0, ss_kz_max=0.0, ss_kz_num=0, mc_kt_num=0, mc_kd_max=0.0,
                      mc_kz_max=0.0, traj_phase_method="CTMC", traj_rtol=None, sample_monte_carlo=False,
                      sample_cutoff_limit=1e-16, rate_prefix="Full", mol_orbit_ridx=0, traj_integrator="tsit5",
                      traj_dt=0.1, precision=None, restep_radius=2.0, seed=0, sample_stratified=False, kd_sigma=0.0,
                      kz_sigma=0.0, reproduce_quirks=True, output_path=None, output_fmt="jld2",
                      output_compress=True, pmd_accumulator="float", pmd_binning="auto", hit_counts=False)


def perform_traj_scan(jobs):
    """A parameter scan on the device: `jobs` is a list of perform_traj_simulation keyword sets, one per scan
    point, as the reference's loops of sequential calls build them (examples/test_Bichromatic_CCP.jl:4-27 over
    intensities, examples/test_Molecules.jl:20-36 over molecules). The points may differ in the laser and in the
    sampler settings (counts, seeds, t0 grid, a molecule's orientation); they must share the target otherwise and
    every setting of the trajectory run (dimension, phase method, integrator, t_final, final momentum grid,
    precision). All points share the device launches (etraj_run_scan); each returns the result dict of
    perform_traj_simulation, equal to what that call alone would return."""
    if not jobs:
        raise ValueError("perform_traj_scan needs at least one job")
    prep = []
    for kw in jobs:
        kw = dict(_SCAN_DEFAULTS, **kw)
        kw["traj_rtol"] = _check_rtol(kw["traj_integrator"], kw["traj_rtol"])
        assert kw["sample_t_intv"][0] < kw["sample_t_intv"][1], \
            "[TrajectorySimulationJob] `sample_t_intv` must be a valid interval."
        assert kw["traj_t_final"] > kw["sample_t_intv"][1], \
            "[TrajectorySimulationJob] `traj_t_final` must be greater than the last time point of `sample_t_intv`."
        assert kw["dimension"] in (2, 3), "[TrajectorySimulationJob] `dimension` must be either 2 or 3."
        assert len(kw["final_p_max"]) == len(kw["final_p_num"]) == kw["dimension"], \
            "[TrajectorySimulationJob] `length(final_p_max)` and `length(final_p_num)` should match `dimension`."
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sp = init_sampler(**kw)
        prep.append((kw, sp, run_desc(**kw)))
    kw0, _, rd0 = prep[0]
    td0 = kw0["target"]._desc()
    for i, (kw, sp, rd) in enumerate(prep):
        if bytes(rd) != bytes(rd0):
            raise ValueError(f"scan job {i}: the trajectory-run settings differ from job 0")
        if bytes(kw["target"]._desc()) != bytes(td0):
            raise ValueError(f"scan job {i}: the target differs from job 0 (only a molecule's orientation may vary)")
    n = len(prep)
    lasers = (_abi.LaserDesc * n)(*[kw["laser"]._desc() for kw, _, _ in prep])
    samplers = (_abi.SamplerDesc * n)(*[sp.desc for _, sp, _ in prep])   # the samplers keep their tables alive
    stats = (_abi.Stats * n)()
    npix = int(np.prod(kw0["final_p_num"]))
    out = np.empty(n * npix)
    hits = np.zeros(n * npix, dtype=np.uint64) if kw0["hit_counts"] else None
    if hits is not None:
        rd0.hit_counts = hits.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64))
    ctx = _abi.default_context()
    ctx.check(ctx.lib.etraj_run_scan(ctx.handle, n, lasers, td0, samplers, rd0, _abi.dptr(out), stats))
    results = []
    for j, (kw, sp, rd) in enumerate(prep):
        st = stats[j]
        r = _result(out[j * npix:(j + 1) * npix].copy(), st.ion_prob, st.ion_prob_uncollected,
                    st.num_effective_traj, st.as_dict(), **kw)
        if hits is not None:
            r["hit_counts"] = _julia_shape(hits[j * npix:(j + 1) * npix].copy(), kw["final_p_num"])
        _save(r, kw["output_path"], kw.get("output_fmt", "jld2"), kw.get("output_compress", True))
        results.append(r)
    return results
