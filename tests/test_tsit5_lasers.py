"""Device Tsit5 through etraj_run_job for the laser forms its kernels treat apart, on the device.

The run_job Tsit5 launch takes a kernel built for a one-colour cos^2/cos^4
laser of at least two cycles with azimuth 0 (phases carried from step to step,
shorter Taylor series for the small phase increments of the stages) and the
generic kernel for every other laser. Each case runs a small ADK grid job (FP64
Tsit5 at traj_rtol 1e-8, CTMC and SCTS) and compares it with the oracle's
run_job: the same electrons, ion probability to 1e-9, step counts to 0.1 %, and
the PMD to a relative L1 of 1e-10 (measured: 1e-15 to 2e-14; an accept/reject
decision that flips on an ulp difference between the two builds moves its
electron by the tolerance).
"""
import math

import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1, small_ctmc
from etraj_hip import _abi, Cos2Laser, Cos4Laser, GaussianLaser, BichromaticLaser

pytestmark = pytest.mark.gpu

W800 = 45.563352525 / 800.0


def _half(cyc):
    return cyc * math.pi / W800


LASERS = {   # name: (laser, half-width of the sampled t0 interval); the first three take the one-colour kernel
    "cos4-lp-2cyc": (Cos4Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0), _half(2)),
    "cos2-cp-3cyc-cep-shift": (Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=3, ellip=1.0, cep=0.7, t_shift=5.0),
                               _half(3)),
    "cos2-2.5cyc-ellip": (Cos2Laser(peak_int=1.2e14, wave_len=800.0, cyc_num=2.5, ellip=0.3), _half(2.5)),
    "cos2-1.5cyc": (Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=1.5, ellip=0.0), _half(1.5)),
    "cos2-azimuth": (Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.5, azi=0.4), _half(2)),
    "gaussian": (GaussianLaser(peak_int=1e14, wave_len=800.0, spread_cyc_num=1.0, ellip=0.0), _half(2)),
    "bichromatic": (BichromaticLaser(Cos2Laser(peak_int=6e13, wave_len=800.0, cyc_num=2, ellip=1.0),
                                     Cos2Laser(peak_int=6e13, wave_len=400.0, cyc_num=4, ellip=-1.0), 0.0), _half(2)),
}


def _device_job(ctx, ld, td, sd, rd):
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    return out, st.as_dict()


@pytest.mark.parametrize("phase", ["CTMC", "SCTS"])
@pytest.mark.parametrize("name", list(LASERS))
def test_tsit5_run_job_laser_forms_match_oracle(ctx, name, phase):
    laser, half = LASERS[name]
    kw = small_ctmc(n_t=80, n_kd=32, traj_integrator="tsit5", traj_rtol=1e-8, precision="fp64",
                    traj_phase_method=phase, laser=laser, sample_t_intv=(-half, half), traj_t_final=half + 10.0)
    ld, td, sd, rd = descriptors(kw)
    got, st = _device_job(ctx, ld, td, sd, rd)
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    err = l1(got, ref)
    print(f"{name} {phase}: {st['num_effective_traj']} electrons, steps {st['n_steps']} (oracle {rst['n_steps']}), "
          f"PMD L1 {err:.1e}", flush=True)
    assert st["num_effective_traj"] == rst["num_effective_traj"] > 100
    assert st["n_failed"] == rst["n_failed"] == 0
    assert st["ion_prob"] == pytest.approx(rst["ion_prob"], rel=1e-9)
    assert abs(st["n_steps"] - rst["n_steps"]) <= 1e-3 * rst["n_steps"]
    assert err <= 1e-10
