#!/usr/bin/env python3
"""Benchmark of the trajectory engine on one GPU.

Workload: H atom (Ip 0.5, Z 1), ADK (:Full prefix), Cos2 800 nm 1e14 W/cm^2 2-cycle
linearly polarized pulse, SCTS phase; 1e4 t0 strata (one seeded jitter each) x
32000 Gaussian k_perp (sigma 0.4 a.u.) per stratum = 3.2e8 samples, of which ~1e8
electrons are above the 1e-16 weight cutoff and integrated; fixed-step RK4
dt = 0.1 a.u. to t_f = 120.3 a.u.; coherent 1000 x 1000 PMD over +-1.5 a.u.
One step = one whole perform_traj_simulation pass (device sampler -> RK4 ->
SCTS phase -> Kepler -> binning -> |amplitude|^2), output left in HBM.

Also measured (outside the timed steps): the same job in all-FP64, the PMD
relative L1 between the mixed and FP64 results, and, on a bounded sample of the
same workload, the CPU FP64 oracle (same RK4 algorithm, and Tsit5 as the
reference's algorithm) timed on the host cores with the GPU's L1 against it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "etraj.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

from etraj_hip import _abi, workloads  # noqa: E402
from etraj_hip.samplers import ADKSampler  # noqa: E402
from etraj_hip.simulation import run_desc  # noqa: E402

# Algorithmic FLOPs of one RK4 step of the 2D SCTS hydrogen-like system (r, p, phase) (add/mul = 1,
# FMA = 2, rsqrt = 1): 4 RHS x 20 + stage states 24 + combination and update 20 + phase 5 = 129,
# in either precision (the FP32 pass's compensation adds and the FP64 folds are implementation,
# not counted). The laser field is read from the per-batch table built by a separate kernel and
# is not counted here.
FLOPS_PER_STEP = {"mixed": 129.0, "fp64": 129.0}
# Algorithmic FLOPs of one attempted Tsit5 step of the same 2D SCTS hydrogen-like system with a
# monochromatic cos^2 LP laser: stage inputs 202, field at 5 stage times by angle addition 545 (+ the
# step's base sin/cos 9), 6 FP64 RHS 108, error sums 70, RMS norm 47, PI controller 14 = 996 (the
# breakdown is in the design notes).
TSIT5_FLOPS_PER_ATTEMPT = 996.0
PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}   # MI355X vector FP32 / FP64 (published)
# The reference's published CPU timings (docs/src/example_overview.md:12-22): ADK-CTMC, 2-cycle CP
# cos^4 4e14 W/cm^2, 54.2M electrons, Tsit5 at traj_rtol 1e-6, AMD Ryzen 9 7950X.
PUBLISHED_CPU = {"us_per_traj_12_cores": 4.3, "us_per_traj_1_core": 23.0, "cpu": "AMD Ryzen 9 7950X",
                 "source": "docs/src/example_overview.md:12-22 (ADK-CTMC 2-cycle CP cos4 4e14, Tsit5 rtol 1e-6)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def descs(kw):
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = ADKSampler(**{k: v for k, v in kw.items() if k != "init_cond_method"})
    return kw["laser"]._desc(), kw["target"]._desc(), sp.desc, run_desc(**kw)


def run_device(ctx, d):
    ptr = _abi.C.POINTER(_abi.C.c_double)()
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job_device(ctx.handle, *d, _abi.C.byref(ptr), _abi.C.byref(st)))
    return st


def run_host(ctx, d, npix):
    out = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, *d, _abi.dptr(out), _abi.C.byref(st)))
    return out, st


def l1(p, q):
    return float(np.abs(p - q).sum() / np.abs(q).sum())


def _device_columns(ctx, ld, td, sd, batches, rows):
    cap = int(ctx.lib.etraj_batch_max_size(sd))
    buf = np.empty((cap, rows))
    cnt = _abi.C.c_int64()
    out = []
    for b in batches:
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b + 1, _abi.dptr(buf), cap,
                                                   _abi.C.byref(cnt)))
        out.append(buf[:cnt.value].copy())
    return np.concatenate(out) if out else np.empty((0, rows))


def _device_records(ctx, ld, td, rd, cols):
    rec = np.empty((len(cols), 12))
    st = np.empty(len(cols), np.int32)
    ctx.check(ctx.lib.etraj_integrate_records(ctx.handle, ld, td, rd, _abi.dptr(np.ascontiguousarray(cols)),
                                              len(cols), cols.shape[1], _abi.dptr(rec),
                                              st.ctypes.data_as(_abi.C.POINTER(_abi.C.c_int32))))
    return rec, st


def _per_electron_errors(rec, st, rrec, rst):
    """final-momentum error |p_inf - p_inf,ref| and phase error mod 2 pi over the electrons that
    both runs count as ionized (status 0)."""
    both = (st == 0) & (rst == 0)
    dp = np.linalg.norm(rec[both, 8:11] - rrec[both, 8:11], axis=1)
    dphi = np.angle(np.exp(1j * (rec[both, 11] - rrec[both, 11])))
    pref = np.linalg.norm(rrec[both, 8:11], axis=1)
    return {"electrons": int(both.sum()), "status_mismatch": int((st != rst).sum()),
            "p_rms": float(np.sqrt(np.mean(dp ** 2))), "p_max": float(dp.max()),
            "p_median": float(np.median(dp)), "p_p99": float(np.quantile(dp, 0.99)),
            "fraction_p_above_1e-6": float(np.mean(dp > 1e-6)),
            "phase_median_mod_2pi": float(np.median(np.abs(dphi))),
            "p_rms_relative": float(np.sqrt(np.mean(dp ** 2)) / np.sqrt(np.mean(pref ** 2))),
            "phase_rms_mod_2pi": float(np.sqrt(np.mean(dphi ** 2))), "phase_max_mod_2pi": float(np.abs(dphi).max())}


def accuracy(kw, ctx, n_strata=64):
    """Per-electron errors of the mixed GPU run against the FP64 GPU run on the same electrons
    (n_strata strata spread over the pulse, sampled on the device)."""
    ldm, tdm, sdm, rdm = descs(kw)
    rd64 = descs(dict(kw, precision="fp64"))[3]
    stride = max(1, sdm.sample_t_num // n_strata)
    rows = (6 if rdm.dimension == 2 else 8) + (1 if rdm.phase_method != _abi.PHASE_CTMC else 0)
    cols = _device_columns(ctx, ldm, tdm, sdm, range(stride // 2, sdm.sample_t_num, stride), rows)
    rec, st = _device_records(ctx, ldm, tdm, rdm, cols)
    r64, s64 = _device_records(ctx, ldm, tdm, rd64, cols)
    out = _per_electron_errors(rec, st, r64, s64)
    out["sample"] = f"{len(cols)} electrons of every {stride}-th stratum"
    return out


def _strata(n_total, n_pick):
    """n_pick strata spread over n_total: the middle one of each of n_pick equal groups."""
    stride = max(1, n_total // n_pick)
    return stride // 2, stride


def host_cpu():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", aff))
    return {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": share, "model": model,
            "threads": max(1, min(aff, share))}


def cpu_timing(kw, threads):
    """The CPU baseline: the oracle (the C FP64 restatement of the reference's CPU method) built
    with -O3 -march=native on this host, batches scheduled dynamically over the threads, timed on
    all the host cores this process may use and on one core, for the reference's algorithm (Tsit5
    at traj_rtol 1e-6) and for the GPU's (RK4 dt 0.1). Each run covers a bounded sample of the
    workload's strata (spread over the pulse) chosen for a few seconds of CPU time."""
    import tempfile
    import oracle as O
    tmp = tempfile.mkdtemp(prefix="etraj_oracle_")
    path = O.build_native(tmp)
    L = O.load_variant(path) if path else None
    build = "gcc -O3 -march=native, built on this host" if L else "prebuilt -O2 (native build failed)"
    out = {"build": build}
    d_ts = descs(dict(kw, precision="fp64", traj_integrator="tsit5", traj_rtol=1e-6))
    d_rk = descs(dict(kw, precision="fp64"))
    n_t = d_ts[2].sample_t_num
    # (threads, strata): the sampled strata are the middle ones of equal groups; the one-core RK4
    # run takes the single stratum at 1/5 of the pulse (the flank of the first field crest)
    plan = {("tsit5", "all"): (threads, 200), ("tsit5", "one"): (1, 12),
            ("rk4", "all"): (threads, 10), ("rk4", "one"): (1, 0)}
    for (alg, cores), (nth, n_strata) in plan.items():
        ld, td, sd, rd = d_ts if alg == "tsit5" else d_rk
        if n_strata:
            first, stride = _strata(n_t, n_strata)
            last = 0
        else:
            first, stride, last = n_t // 5, 1, n_t // 5 + 1
        t = time.perf_counter()
        _, st = O.run_job(ld, td, sd, rd, nthreads=nth, batch_begin=first, batch_end=last, batch_stride=stride,
                          dynamic=True, library=L)
        dt = time.perf_counter() - t
        n = st["num_effective_traj"]
        out[f"{alg}_{'all_cores' if cores == 'all' else '1_core'}"] = {"value": n / dt, "unit": "trajectories/s", "threads": nth, "seconds": dt,
                                       "electrons": n, "us_per_traj": 1e6 * dt / n,
                                       "us_per_traj_x_cores": 1e6 * dt * nth / n,
                                       "sample": (f"every {stride}-th of {n_t} strata from {first}" if n_strata
                                                  else f"stratum {first} of {n_t}")}
        log(f"cpu {alg} on {nth} threads: {n / dt:.0f} traj/s ({n} electrons, {dt:.1f} s)")
    return out


def reference_comparison(kw, ctx, n_strata, threads):
    """The GPU headline against the reference's own algorithm on the same electrons: on n_strata
    strata spread over the pulse, the oracle's FP64 Tsit5 at traj_rtol 1e-6 (the reference's
    integrator, TrajectorySimulation.jl:342) and FP64 RK4 dt 0.1 PMDs, against the device's mixed
    RK4 (the headline's algorithm) and FP64 Tsit5 on the same columns (gen_electron_batch +
    integrate_and_bin); per-electron momentum and phase errors of the headline against the
    oracle's Tsit5 and RK4 on every 4th electron of 8 of the strata."""
    import oracle as O
    d64 = descs(dict(kw, precision="fp64"))
    dts = descs(dict(kw, precision="fp64", traj_integrator="tsit5", traj_rtol=1e-6))
    dmx = descs(kw)
    dgts = descs(dict(kw, precision="fp64", traj_integrator="tsit5", traj_rtol=1e-6))
    ld, td, sd, _ = d64
    first, stride = _strata(sd.sample_t_num, n_strata)
    ref_ts, st_ts = O.run_job(*dts, nthreads=threads, batch_begin=first, batch_stride=stride)
    ref_rk, _ = O.run_job(*d64, nthreads=threads, batch_begin=first, batch_stride=stride)
    npix = dmx[3].final_p_num[0] * dmx[3].final_p_num[1]
    rows = 7
    cap = int(ctx.lib.etraj_batch_max_size(sd))
    buf = np.empty((cap, rows))
    cnt = _abi.C.c_int64()
    amp = {"mixed_rk4": np.zeros(2 * npix), "fp64_tsit5": np.zeros(2 * npix)}
    for b in range(first, sd.sample_t_num, stride):
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b + 1, _abi.dptr(buf), cap,
                                                   _abi.C.byref(cnt)))
        if not cnt.value:
            continue
        for key, rd in (("mixed_rk4", dmx[3]), ("fp64_tsit5", dgts[3])):
            ip, iu, nf = _abi.C.c_double(0), _abi.C.c_double(0), _abi.C.c_int64(0)
            ctx.check(ctx.lib.etraj_integrate_and_bin(ctx.handle, ld, td, rd, _abi.dptr(buf), cnt.value, rows,
                                                      amp[key].ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip),
                                                      _abi.C.byref(iu), _abi.C.byref(nf)))
    gpu = {k: v[0::2] ** 2 + v[1::2] ** 2 for k, v in amp.items()}
    cols = _device_columns(ctx, ld, td, sd, range(first, sd.sample_t_num, stride * max(1, n_strata // 8)), rows)[::4]
    rts, sts = O.integrate_records(*dts[:2], dts[3], cols)
    rrk, srk = O.integrate_records(ld, td, d64[3], cols)
    rmx, smx = _device_records(ctx, ld, td, dmx[3], cols)
    r64, s64 = _device_records(ctx, ld, td, d64[3], cols)
    return {
        "sample": f"{st_ts['num_effective_traj']} electrons = every {stride}-th of {sd.sample_t_num} strata from {first}",
        "l1_vs_reference_algorithm": l1(gpu["mixed_rk4"], ref_ts),
        "l1_oracle_rk4_vs_oracle_tsit5": l1(ref_rk, ref_ts),
        "l1_gpu_mixed_rk4_vs_oracle_rk4": l1(gpu["mixed_rk4"], ref_rk),
        "l1_gpu_tsit5_vs_oracle_tsit5": l1(gpu["fp64_tsit5"], ref_ts),
        "per_electron": {"sample": f"{len(cols)} electrons",
                         "gpu_mixed_rk4_vs_oracle_tsit5": _per_electron_errors(rmx, smx, rts, sts),
                         "gpu_mixed_rk4_vs_oracle_rk4": _per_electron_errors(rmx, smx, rrk, srk),
                         "gpu_fp64_rk4_vs_oracle_rk4": _per_electron_errors(r64, s64, rrk, srk)},
    }


def tsit5_entry(kw, ctx, steps, warmup):
    """The headline workload through the device's adaptive Tsit5 at traj_rtol 1e-6 in FP64 (the
    reference's integrator): whole-job throughput and the k_tsit5 roofline."""
    d = descs(dict(kw, precision="fp64", traj_integrator="tsit5", traj_rtol=1e-6))
    for _ in range(warmup):
        run_device(ctx, d)
    t = time.perf_counter()
    sts = [run_device(ctx, d) for _ in range(steps)]
    sec = (time.perf_counter() - t) / steps
    st = sts[-1]
    launch_ms = float(np.mean([s.ms_integrate_main for s in sts])) / max(st.n_slabs, 1)
    f = TSIT5_FLOPS_PER_ATTEMPT * st.n_attempts / max(st.n_slabs, 1)
    ach = f / (launch_ms / 1e3) / 1e12
    return {"value": st.num_effective_traj / sec, "unit": "trajectories/s", "ms_per_step": sec * 1e3,
            "electrons": st.num_effective_traj, "steps_accepted": st.n_steps, "steps_attempted": st.n_attempts,
            "failed": st.n_failed, "dtype": "fp64",
            "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                         "frac": ach / PEAK_TFLOPS["fp64"], "kernel": "k_tsit5 FP64",
                         "launches_per_job": st.n_slabs, "avg_launch_ms": launch_ms, "flops_per_launch": f,
                         "flops_per_attempt": TSIT5_FLOPS_PER_ATTEMPT, "timing": "HIP events on the ctx stream"}}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_rk4_h_lp_scts_strat_1e8_pmc.md")


def traffic_per_launch():
    """HBM bytes per launch of the FP32 pass, from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (the json block of the profile file), or None if that file is absent."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)["traffic_bytes_per_launch"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


PROBE = os.path.join(ROOT, "etraj.jl_amd", "etraj_hip", "valu_probe")


def valu_probe():
    """The vector-ALU rates this device reaches with non-packed FP32 and FP64 instructions (valu_probe:
    16 independent FMA / add / rsq chains per lane, vector operands, 8 waves per SIMD), or None if the probe
    executable is missing. The published FP32 peak counts packed FP32 FMAs, which this build never emits."""
    import subprocess
    if not os.path.exists(PROBE):
        return None
    r = subprocess.run([PROBE], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        log(f"valu_probe failed: {r.stderr.strip()}")
        return None
    return json.loads(r.stdout)


def copy_calibration(mib):
    """One hipMemcpy device-to-device of `mib` MiB: a copy kernel of known traffic (mib MiB read +
    mib MiB written) whose FETCH_SIZE / WRITE_SIZE counters scale the integrator's."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0          # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {mib} MiB device-to-device")



# ---- the other configurations (python bench.py --workload NAME) -------------------

def _scan_descs(jobs):
    import warnings
    from etraj_hip.samplers import init_sampler
    ds, keep = [], []
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for kw in jobs:
            sp = init_sampler(**kw)
            keep.append(sp)
            ds.append((kw["laser"]._desc(), kw["target"]._desc(), sp.desc, run_desc(**kw)))
    return ds, keep


def _run_scan(ctx, ds, device=True):
    n = len(ds)
    ld = (_abi.LaserDesc * n)(*[d[0] for d in ds])
    sd = (_abi.SamplerDesc * n)(*[d[2] for d in ds])
    st = (_abi.Stats * n)()
    rd = ds[0][3]
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    if device:
        ptr = _abi.C.POINTER(_abi.C.c_double)()
        ctx.check(ctx.lib.etraj_run_scan_device(ctx.handle, n, ld, ds[0][1], sd, rd, _abi.C.byref(ptr), st))
        return None, [st[j] for j in range(n)]
    out = np.empty(n * npix)
    ctx.check(ctx.lib.etraj_run_scan(ctx.handle, n, ld, ds[0][1], sd, rd, _abi.dptr(out), st))
    return out.reshape(n, npix), [st[j] for j in range(n)]


def _timed(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    sts = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sts.append(fn())
    return (time.perf_counter() - t0) / steps, sts


def _roofline_rk4(st_list, steps, flops32, flops64):
    """Roofline of the FP32 pass from summed per-call stats (RK4 mixed runs)."""
    s0 = st_list[0]
    if not s0.ms_integrate_main > 0:   # graph replay: no per-kernel timing
        return None
    n_steps = sum(s.n_steps for s in st_list) / steps
    n64 = sum(s.n_steps_fp64 for s in st_list) / steps
    launch_ms = s0.ms_integrate_main / max(s0.n_slabs, 1)
    f = flops32 * (n_steps - n64) / max(s0.n_slabs, 1)
    ach = f / (launch_ms / 1e3) / 1e12
    return {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s",
            "frac": ach / PEAK_TFLOPS["mixed"], "traffic": None, "kernel": "k_rk4 FP32 pass",
            "launches_per_job": s0.n_slabs, "avg_launch_ms": launch_ms, "flops_per_launch": f,
            "flops_per_step": {"fp32_pass": flops32, "fp64": flops64}}


def _cpu_sample(ds, threads, n_jobs_sample=2, stride=None):
    """The oracle (-O3 -march=native build, dynamic schedule) on a bounded sample: every stride-th
    batch of the first n_jobs_sample jobs."""
    import tempfile
    import oracle as O
    path = O.build_native(tempfile.mkdtemp(prefix="etraj_oracle_"))
    L = O.load_variant(path) if path else None
    n = 0
    t = time.perf_counter()
    for (ld, td, sd, rd) in ds[:n_jobs_sample]:
        s = stride or max(1, sd.sample_t_num // 8)
        _, st = O.run_job(ld, td, sd, rd, nthreads=threads, batch_begin=s // 2, batch_stride=s, dynamic=True,
                          library=L)
        n += st["num_effective_traj"]
    return n, time.perf_counter() - t


def bench_workload(name, args, ctx):
    threads = host_cpu()["threads"]
    if name == "h-ctmc-grid":
        kw = workloads.h_lp_ctmc_grid(n_t=2000, n_kd=50)
        jobs = [dict(kw, precision="mixed", pmd_accumulator="fixed")]
        desc = ("H-atom ADK CTMC, 800 nm 1e14 W/cm^2 2-cycle LP cos2, 2000 t0 x 50 kd grid, RK4 dt 0.1, "
                "PMD 500x500: fixed-point weight sums + per-cell hit counts")
    elif name == "h-scts-1e8-tsit5":
        kw = dict(workloads.h_lp_scts_stratified(n_t=args.n_t or 10000, n_k=args.n_k or 34400, precision="fp64"),
                  traj_integrator="tsit5", traj_rtol=1e-6)
        jobs = [kw]
        desc = ("H-atom ADK+SCTS, 800 nm 1e14 W/cm^2 2-cycle LP cos2, "
                f"{kw['sample_t_num']} t0 strata x {kw['mc_kt_num']} Gaussian k_perp samples, adaptive Tsit5 "
                "traj_rtol 1e-6 in FP64 (the reference's integrator), coherent PMD 1000x1000")
    elif name == "ar-spa-3d":
        kw = workloads.ar_cp_spa_qtmc_3d(n_t=args.n_t or 2000, n_k=args.n_k or 158)
        jobs = [kw]
        desc = (f"Ar (SAE, l=0 prefactor) SFA-SPA 3D grid {kw['sample_t_num']} x {kw['ss_kd_num']} x "
                f"{kw['ss_kz_num']}, CP 800 nm 2e14 W/cm^2 2 cycles, Tsit5 rtol 1e-8 ({kw['precision']}), QTMC, "
                "PMD 256^3")
    elif name == "bichromatic-scan":
        ints = list(np.logspace(np.log10(5e13), np.log10(5e14), 32))
        probe, _ = _scan_descs(workloads.bichromatic_scan_jobs(ints, n_t=200, n_k=200))
        _, pst = _run_scan(ctx, probe)
        n_k = workloads.bichromatic_scan_counts([s.ion_prob for s in pst], [s.num_effective_traj for s in pst],
                                                200 * 200)
        jobs = workloads.bichromatic_scan_jobs(ints, n_t=2000, n_k=n_k)
        desc = ("32 intensities 5e13-5e14 W/cm^2, bichromatic counter-rotating CP 800+400 nm 8/16 cycles, "
                "H ADK+SCTS, counts per intensity proportional to the ADK yield (1e5-2e7), 2D PMD 512x512 each")
    elif name == "wfat-orientation-scan":
        jobs = workloads.wfat_orientation_jobs(args.scan_jobs or 64, n_t=2000, n_k=2000)
        desc = (f"{len(jobs)} orientations (seeded uniform SO(3)), single-centre Z=1 Ip=0.45 WFAT G00 8x16 table, "
                "1300 nm 4-cycle LP 1e14 W/cm^2, SCTS, 2D PMD 400x400 each")
    else:
        raise SystemExit(f"unknown workload {name}")
    jobs = [dict(j, pmd_binning=args.pmd_binning) for j in jobs]
    tuned = None
    if args.autotune and len(jobs) == 1:
        from etraj_hip import autotune
        tuned = autotune(**jobs[0])
        log(f"autotune: {tuned['start_ms']:.1f} ms -> {tuned['best_ms']:.1f} ms with {tuned['best']}")
    ds, keep = _scan_descs(jobs)
    hits = None
    if name == "h-ctmc-grid":   # the integer-count PMD: electrons binned per cell, accumulated over the runs
        hits = np.zeros(int(np.prod(jobs[0]["final_p_num"])), dtype=np.uint64)
        ds[0][3].hit_counts = hits.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64))
    rk4 = ds[0][3].integrator == _abi.INTEG_RK4
    mixed = ds[0][3].precision == _abi.PREC_MIXED
    # a job of a few ms is timed over enough calls that the timed region lasts about 0.1 s
    steps, warmup = args.steps, args.warmup
    if name == "h-ctmc-grid":
        steps, warmup = max(steps, 60), max(warmup, 10)
    sec, sts = _timed(ctx, lambda: _run_scan(ctx, ds)[1], steps, warmup)
    n_el = sum(s.num_effective_traj for s in sts[-1])
    result = {
        "metric": "trajectories/s + PMD wall time, mixed-prec vs FP64 GPU at PMD L1<=1e-3",
        "value": n_el / sec, "unit": "trajectories/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64" if mixed else "fp64", "data": "synthetic (seeded sampling)",
        "config": {"workload": desc, "electrons": n_el, "jobs": len(ds)},
        "pmd_wall_ms": sec * 1e3,
        "tuning": ctx.get_tuning(),
        "autotune": tuned,
        "device_ms": {"sample": sts[-1][0].ms_sample, "integrate": sts[-1][0].ms_integrate,
                      "integrate_main": sts[-1][0].ms_integrate_main},
        "fp64_reintegrated_fraction": sum(s.n_restep for s in sts[-1]) / max(n_el, 1),
        "pmd_accumulator": "fixed" if ds[0][3].pmd_mode == _abi.PMD_FIXED else "float",
        # algorithmic FLOPs per RK4 step: 2D SCTS hydrogen-like 129 (+8 for the molecule's soft-core
        # term, 2 per RHS); 2D CTMC 100 (no phase: 4 RHS x 6 + phase combination 5 fewer)
        "roofline": (_roofline_rk4(sts[-1], 1, {"wfat-orientation-scan": 137.0, "h-ctmc-grid": 100.0}.get(name, 129.0),
                                   129.0) if (mixed and rk4) else None),
    }
    if name == "h-scts-1e8-tsit5":
        st = sts[-1][0]
        launch_ms = st.ms_integrate_main / max(st.n_slabs, 1)
        f = TSIT5_FLOPS_PER_ATTEMPT * st.n_attempts / max(st.n_slabs, 1)
        ach = f / (launch_ms / 1e3) / 1e12
        result["roofline"] = {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                              "frac": ach / PEAK_TFLOPS["fp64"], "traffic": None, "kernel": "k_tsit5 FP64",
                              "launches_per_job": st.n_slabs, "avg_launch_ms": launch_ms, "flops_per_launch": f,
                              "flops_per_attempt": TSIT5_FLOPS_PER_ATTEMPT, "steps_attempted": st.n_attempts,
                              "steps_accepted": st.n_steps}
    if hits is not None:
        result["hit_counts_per_run"] = int(hits.sum()) // (steps + warmup)
        ds[0][3].hit_counts = None
    ds64 = None
    if mixed and not args.no_fp64:
        ds64, keep64 = _scan_descs([dict(kw, precision="fp64") for kw in jobs])
        sec64, _ = _timed(ctx, lambda: _run_scan(ctx, ds64)[1], steps, warmup)
        p64, s64 = _run_scan(ctx, ds64, device=False)
        pmx, smx = _run_scan(ctx, ds, device=False)
        result["fp64_gpu"] = {"value": sum(s.num_effective_traj for s in s64) / sec64, "pmd_wall_ms": sec64 * 1e3}
        result["mixed_over_fp64"] = sec64 / sec
        result["pmd_l1_mixed_vs_fp64_gpu"] = max(l1(pmx[j], p64[j]) for j in range(len(ds)))
    if args.cpu_batches > 0:
        if mixed and ds64 is None:
            ds64, keep64 = _scan_descs([dict(kw, precision="fp64") for kw in jobs])
        cds = [(ld, td, sd, rd) for (ld, td, sd, rd) in (ds64 if mixed else ds)]
        n, dt = _cpu_sample(cds, threads, n_jobs_sample=min(2, len(cds)),
                            stride=max(1, cds[0][2].sample_t_num // args.cpu_batches))
        result["cpu_baseline"] = {"value": n / dt, "unit": "trajectories/s", "cores": threads, "kind": "port",
                                  "sample": f"{n} electrons, every {max(1, cds[0][2].sample_t_num // args.cpu_batches)}"
                                            f"-th batch of {min(2, len(cds))} job(s), the oracle's "
                                            f"{'Tsit5' if not rk4 else 'FP64 RK4'}, {dt:.1f} s"}
    return result

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-t", type=int, default=None, help="t0 samples (default: the workload's own)")
    ap.add_argument("--n-k", type=int, default=None, help="momentum samples per t0 (default: the workload's own)")
    ap.add_argument("--cpu-batches", type=int, default=24, help="strata in the CPU baseline sample (0: skip)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tsit5", action="store_true", help="skip the device Tsit5 entry")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (HBM-counter calibration)")
    ap.add_argument("--workload", default="h-scts-1e8",
                    choices=["h-scts-1e8", "h-scts-1e8-tsit5", "h-ctmc-grid", "ar-spa-3d", "bichromatic-scan",
                             "wfat-orientation-scan"],
                    help="h-scts-1e8 (default) is the headline; the others are the remaining configurations")
    ap.add_argument("--scan-jobs", type=int, default=0, help="orientations of wfat-orientation-scan (default 64)")
    ap.add_argument("--restep-radius", type=float, default=None,
                    help="FP64 re-integration radius of the mixed mode (a.u.; default: the workload's)")
    ap.add_argument("--pmd-binning", default="auto", choices=["auto", "atomic", "tiles"],
                    help="PMD binning: the context's choice, per-electron atomics, or LDS-privatised tiles")
    ap.add_argument("--tuning", default="", metavar="K=V,...",
                    help="launch choices set before the run (etraj_tuning fields, e.g. rk4_block=128,bin_tiles=1)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the engine's autotuner on the workload first and time the chosen launch set")
    ap.add_argument("--no-valu-probe", action="store_true",
                    help="skip the vector-ALU throughput probe (the non-packed FP32 / FP64 ceiling of the roofline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the trajectory path shards by "
              "t0 strata with one PMD reduction, which this build does not run across GPUs.", file=sys.stderr)
        sys.exit(2)

    probe = None if args.no_valu_probe else valu_probe()
    ctx = _abi.default_context()
    if args.tuning:
        ctx.set_tuning(**{k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))})
    if args.calibrate_copy > 0:
        copy_calibration(args.calibrate_copy)
    if args.workload != "h-scts-1e8":
        res = bench_workload(args.workload, args, ctx)
        if probe:
            res["valu_probe"] = probe
            rf = res.get("roofline")
            if rf:
                key = "fma_f64" if rf["peak"] == PEAK_TFLOPS["fp64"] else "fma_f32"
                rf[f"probe_{key}_tflops"] = probe[key]["tflops"]
                rf[f"frac_of_probe_{key}"] = rf["achieved"] / probe[key]["tflops"]
        print(json.dumps(res), flush=True)
        return
    kw = workloads.h_lp_scts_stratified(n_t=args.n_t or 10000, n_k=args.n_k or 34400, precision="mixed")
    if args.restep_radius is not None:
        kw["restep_radius"] = args.restep_radius
    kw["pmd_binning"] = args.pmd_binning
    tuned = None
    if args.autotune:
        from etraj_hip import autotune
        tuned = autotune(**kw)
        log(f"autotune: {tuned['start_ms']:.1f} ms -> {tuned['best_ms']:.1f} ms with {tuned['best']}")
    d_mixed = descs(kw)
    npix = kw["final_p_num"][0] * kw["final_p_num"][1]

    for i in range(args.warmup):
        st = run_device(ctx, d_mixed)
        log(f"warmup {i}: {st.ms_total:.1f} ms, {st.num_effective_traj} electrons")
    ms, ms_int, ms_main, steps, steps64, restep, n_el, n_slabs = [], [], [], 0, 0, 0, 0, 1
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = run_device(ctx, d_mixed)
        ms.append(st.ms_total)
        ms_int.append(st.ms_integrate)
        ms_main.append(st.ms_integrate_main)
        n_slabs = st.n_slabs
        steps += st.n_steps
        steps64 += st.n_steps_fp64
        restep += st.n_restep
        n_el = st.num_effective_traj
        log(f"step {i}: {st.ms_total:.1f} ms (integrate kernels {st.ms_integrate:.1f} ms), "
            f"{st.n_restep} electrons re-integrated in FP64")
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = n_el / (ms_step / 1e3)

    # Dominant kernel: the FP32 pass (every electron, one launch per slab), timed with HIP events on the
    # ctx stream it runs on. The FP64 pass re-integrates the electrons that came within restep_radius.
    flops32 = FLOPS_PER_STEP["mixed"] * (steps - steps64) / args.steps
    flops_all = flops32 + FLOPS_PER_STEP["fp64"] * steps64 / args.steps
    launch_ms = float(np.mean(ms_main)) / n_slabs
    achieved = flops32 / n_slabs / (launch_ms / 1e3) / 1e12
    result = {
        "metric": "trajectories/s + PMD wall time, mixed-prec vs FP64 GPU at PMD L1<=1e-3",
        "value": value, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic (seeded Philox sampling of the ADK distribution)",
        "config": {"workload": "H-atom ADK+SCTS, 800 nm 1e14 W/cm^2 2-cycle LP cos2, "
                               f"{kw['sample_t_num']} t0 strata x {kw['mc_kt_num']} Gaussian k_perp samples "
                               f"({n_el} electrons above the 1e-16 cutoff), RK4 dt 0.1, coherent PMD 1000x1000",
                   "electrons": n_el,
                   "precision": "mixed: FP32 EOM + Kahan, FP32/FP64 Kahan phase, FP64 re-integration of electrons "
                                f"within {kw['restep_radius']:g} a.u. of the core, FP64 Kepler and PMD"},
        "pmd_wall_ms": ms_step,
        "tuning": ctx.get_tuning(),
        "autotune": tuned,
        "integrate_kernel_ms": float(np.mean(ms_int)),
        "fp64_reintegrated_fraction": restep / max(n_el * args.steps, 1),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["mixed"], "traffic": traffic_per_launch(),
                     "kernel": "k_rk4 FP32 pass", "launches_per_job": n_slabs, "avg_launch_ms": launch_ms,
                     "flops_per_launch": flops32 / n_slabs,
                     "flops_per_step": {"fp32_pass": FLOPS_PER_STEP["mixed"], "fp64": FLOPS_PER_STEP["fp64"]},
                     "rk4_steps_per_job": steps // args.steps, "fp64_steps_per_job": steps64 // args.steps,
                     "both_passes_tflops": flops_all / (np.mean(ms_int) / 1e3) / 1e12},
    }
    if probe:
        result["valu_probe"] = probe
        ceil = probe["fma_f32"]["tflops"]
        result["roofline"]["probe_fma_f32_tflops"] = ceil
        result["roofline"]["frac_of_probe_fma_f32"] = achieved / ceil
    if not args.no_fp64:
        d64 = descs(dict(kw, precision="fp64"))
        p64, s64 = run_host(ctx, d64, npix)
        pmx, smx = run_host(ctx, d_mixed, npix)
        result["fp64_gpu"] = {"value": s64.num_effective_traj / (s64.ms_total / 1e3), "pmd_wall_ms": s64.ms_total,
                              "integrate_kernel_ms": s64.ms_integrate,
                              "achieved_tflops": FLOPS_PER_STEP["fp64"] * s64.n_steps / (s64.ms_integrate / 1e3) / 1e12,
                              "frac_of_fp64_peak": FLOPS_PER_STEP["fp64"] * s64.n_steps / (s64.ms_integrate / 1e3)
                              / 1e12 / PEAK_TFLOPS["fp64"]}
        result["mixed_over_fp64"] = (smx.num_effective_traj / smx.ms_total) / (s64.num_effective_traj / s64.ms_total)
        # etraj_run_job with the PMD returned in a host buffer (the drop-in's host hand-over)
        result["host_output_value"] = smx.num_effective_traj / (smx.ms_total / 1e3)
        result["pmd_l1_mixed_vs_fp64_gpu"] = l1(pmx, p64)
        result["ion_prob"] = {"mixed": smx.ion_prob, "fp64": s64.ion_prob}
        log(f"fp64: {s64.ms_total:.1f} ms; L1(mixed, fp64) = {result['pmd_l1_mixed_vs_fp64_gpu']:.3e}")
    if not args.no_fp64:
        result["accuracy_mixed_vs_fp64_gpu"] = accuracy(kw, ctx)
        a = result["accuracy_mixed_vs_fp64_gpu"]
        log(f"per electron (mixed vs fp64 GPU, {a['electrons']}): p_inf RMS {a['p_rms']:.2e}, "
            f"phase RMS mod 2pi {a['phase_rms_mod_2pi']:.2e}")
    if not args.no_tsit5:
        result["tsit5_entry"] = tsit5_entry(kw, ctx, max(1, args.steps // 2), 1)
        if probe:
            rf = result["tsit5_entry"]["roofline"]
            rf["probe_fma_f64_tflops"] = probe["fma_f64"]["tflops"]
            rf["frac_of_probe_fma_f64"] = rf["achieved"] / probe["fma_f64"]["tflops"]
        log(f"tsit5 entry: {result['tsit5_entry']['ms_per_step']:.1f} ms per job")
    if args.cpu_batches > 0:
        cpu = host_cpu()
        result["reference_algorithm"] = reference_comparison(kw, ctx, args.cpu_batches, cpu["threads"])
        log(f"L1 vs the reference algorithm: {result['reference_algorithm']['l1_vs_reference_algorithm']:.3e}")
        tm = cpu_timing(kw, cpu["threads"])
        ta = tm["tsit5_all_cores"]
        result["cpu_baseline"] = {
            "value": ta["value"], "unit": "trajectories/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"the oracle's FP64 Tsit5 at traj_rtol 1e-6 (the reference's algorithm), {tm['build']}, "
                      f"{ta['electrons']} electrons ({ta['sample']}), {ta['seconds']:.1f} s on {cpu['threads']} threads",
            "host": cpu, "timings": tm, "published": PUBLISHED_CPU}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
