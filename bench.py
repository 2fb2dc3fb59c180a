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

# Algorithmic FLOPs of one RK4 step of the 2D SCTS hydrogen-like integrator kernel (add/mul = 1,
# FMA = 2, rsqrt = 1): 4 RHS x 20 + stage states 24 + combination 20 + phase 5 = 129, plus the
# state update (FP32 pass: 4 x FP32 Kahan 4 + FP32 Kahan phase 4 = 20; FP64: 5). The laser field
# is read from the per-batch table built by a separate kernel and is not counted here.
FLOPS_PER_STEP = {"mixed": 149.0, "fp64": 134.0}
PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}   # MI355X vector FP32 / FP64 (published)


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


def cpu_baseline(kw, ctx, n_batches_sample, threads):
    """Oracle (FP64 RK4, same algorithm; and Tsit5 rtol 1e-6, the reference's algorithm) on
    n_batches_sample strata spread evenly over the pulse (the middle stratum of each of
    n_batches_sample equal groups); the GPU runs the same strata through gen_electron_batch +
    integrate_and_bin for the L1 comparison."""
    import oracle as O
    ld, td, sd, rd = descs(dict(kw, precision="fp64"))
    stride = max(1, sd.sample_t_num // n_batches_sample)
    first = stride // 2
    t = time.perf_counter()
    ref, st = O.run_job(ld, td, sd, rd, nthreads=threads, batch_begin=first, batch_stride=stride)
    dt_rk4 = time.perf_counter() - t
    n = st["num_effective_traj"]
    # reference algorithm on every second of those strata
    ldt, tdt, sdt, rdt = descs(dict(kw, precision="fp64", traj_integrator="tsit5"))
    t = time.perf_counter()
    _, stt = O.run_job(ldt, tdt, sdt, rdt, nthreads=threads, batch_begin=first, batch_stride=2 * stride)
    dt_ts = time.perf_counter() - t
    # the GPU (mixed) on the same strata
    ldm, tdm, sdm, rdm = descs(kw)
    npix = rdm.final_p_num[0] * rdm.final_p_num[1]
    amp = np.zeros(2 * npix)
    ip, iu, nf = _abi.C.c_double(0), _abi.C.c_double(0), _abi.C.c_int64(0)
    cap = int(ctx.lib.etraj_batch_max_size(sdm))
    buf = np.empty((cap, 7))
    cnt = _abi.C.c_int64()
    for b in range(first, sdm.sample_t_num, stride):
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ldm, tdm, sdm, b + 1, _abi.dptr(buf), cap,
                                                   _abi.C.byref(cnt)))
        if cnt.value:
            ctx.check(ctx.lib.etraj_integrate_and_bin(ctx.handle, ldm, tdm, rdm, _abi.dptr(buf), cnt.value, 7,
                                                      amp.ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip),
                                                      _abi.C.byref(iu), _abi.C.byref(nf)))
    gpu = amp[0::2] ** 2 + amp[1::2] ** 2
    return {
        "value": n / dt_rk4, "unit": "trajectories/s", "cores": threads, "kind": "port",
        "sample": f"{n} electrons = the strata {first}, {first + stride}, ... (every {stride}-th of "
                  f"{sd.sample_t_num}), FP64 RK4 dt 0.1 (same algorithm as the GPU), {dt_rk4:.1f} s",
        "reference_algorithm_value": stt["num_effective_traj"] / dt_ts,
        "reference_algorithm_sample": f"{stt['num_effective_traj']} electrons (every {2 * stride}-th stratum), "
                                      f"Tsit5 rtol 1e-6 (the reference's integrator), {dt_ts:.1f} s",
        "gpu_mixed_pmd_l1_vs_cpu_fp64": l1(gpu, ref),
    }


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-t", type=int, default=10000)
    ap.add_argument("--n-k", type=int, default=32000)
    ap.add_argument("--cpu-batches", type=int, default=24, help="strata in the CPU baseline sample (0: skip)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the trajectory path shards by "
              "t0 strata with one PMD reduction, which this build does not run across GPUs.", file=sys.stderr)
        sys.exit(2)

    ctx = _abi.default_context()
    if args.calibrate_copy > 0:
        copy_calibration(args.calibrate_copy)
    kw = workloads.h_lp_scts_stratified(n_t=args.n_t, n_k=args.n_k, precision="mixed")
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
                                "within 2 a.u. of the core, FP64 Kepler and PMD"},
        "pmd_wall_ms": ms_step,
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
    if args.cpu_batches > 0:
        threads = int(os.environ.get("ETRAJ_CPU_THREADS", min(os.cpu_count() or 1, 16)))
        result["cpu_baseline"] = cpu_baseline(kw, ctx, args.cpu_batches, threads)
        log("cpu baseline:", result["cpu_baseline"]["sample"])
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
