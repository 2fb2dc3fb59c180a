"""Writes the golden fixtures under tests/golden/ from the CPU FP64 oracle.

    python tests/golden/make_golden.py

Fixtures (sparse PMDs keep them small):
  kat_counts.json          the reference's sampler survivor counts (ElectronSamplers_test.jl:53-128)
  ctmc_grid_small.npz      H, ADK, LP 1e14 800 nm 2 cyc, 100 t0 x 32 kd, RK4 dt 0.1, CTMC 500x500
  scts_strat_small.npz     same pulse, SCTS, 40 strata x 256 Gaussian k (sigma 0.4, seed 0xE7A10002), 200x200
                           amplitude sum (re, im) before abs2
  records_small.npz        per-electron records (12 doubles + status) of 13 batches of the CTMC grid
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "etraj.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import oracle as O  # noqa: E402
from cases import descriptors, small_ctmc, small_scts  # noqa: E402

KATS = [
    ({}, [6, 590]), ({"traj_phase_method": "QTMC"}, [7, 590]), ({"traj_phase_method": "SCTS"}, [7, 590]),
    ({"dimension": 3, "ss_kd_num": 30, "ss_kz_num": 30}, [8, 376]),
    ({"dimension": 3, "ss_kd_num": 30, "ss_kz_num": 30, "traj_phase_method": "QTMC"}, [9, 376]),
    ({"dimension": 3, "ss_kd_num": 30, "ss_kz_num": 30, "traj_phase_method": "SCTS"}, [9, 376]),
    ({"rate_prefix": "Exp"}, [6, 536]), ({"rate_prefix": "Pre"}, [6, 562]), ({"rate_prefix": "PreCC"}, [6, 610]),
    ({"rate_prefix": "Jac"}, [6, 514]), ({"rate_prefix": ["Pre", "Jac"]}, [6, 540]),
]


def sparse(a):
    idx = np.flatnonzero(a)
    return idx.astype(np.int64), a[idx]


def main():
    with open(os.path.join(HERE, "kat_counts.json"), "w") as f:
        json.dump({"source": "ElectronSamplers_test.jl:53-128, size(gen_electron_batch(sp, 200))",
                   "base": {"laser": "Cos4Laser(peak_int=1e15, wave_len=800, cyc_num=2, ellip=1)", "target": "H",
                            "sample_t_intv": [-100, 100], "sample_t_num": 400, "sample_cutoff_limit": 1e-14,
                            "ss_kd_max": 2.0, "ss_kd_num": 800, "ss_kz_max": 2.0},
                   "cases": [{"override": o, "size": s} for o, s in KATS]}, f, indent=1)

    kw = small_ctmc(n_t=100, n_kd=32)
    ld, td, sd, rd = descriptors(kw)
    pmd, st = O.run_job(ld, td, sd, rd, nthreads=1)
    i, v = sparse(pmd)
    np.savez_compressed(os.path.join(HERE, "ctmc_grid_small.npz"), idx=i, val=v, npix=pmd.size,
                        ion_prob=st["ion_prob"], ion_prob_uncollected=st["ion_prob_uncollected"],
                        num_effective_traj=st["num_effective_traj"])

    kw = small_scts(n_t=40, n_k=256)
    ld, td, sd, rd = descriptors(kw)
    amp, st = O.run_job_amplitude(ld, td, sd, rd)
    i, v = sparse(amp)
    np.savez_compressed(os.path.join(HERE, "scts_strat_small.npz"), idx=i, val=v, n=amp.size,
                        ion_prob=st["ion_prob"], num_effective_traj=st["num_effective_traj"])

    kw = small_ctmc(n_t=60, n_kd=40)
    ld, td, sd, rd = descriptors(kw)
    cols = np.concatenate([c for c in (O.gen_electron_batch(ld, td, sd, b) for b in range(8, 60, 4)) if c is not None])
    rec, status = O.integrate_records(ld, td, rd, cols)
    np.savez_compressed(os.path.join(HERE, "records_small.npz"), cols=cols, rec=rec, status=status)
    print("written to", HERE)


if __name__ == "__main__":
    main()
