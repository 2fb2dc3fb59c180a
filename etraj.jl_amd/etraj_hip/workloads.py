"""Named workloads of the engine (keyword sets for perform_traj_simulation).

h_lp_ctmc_grid        H atom (Ip 0.5, Z 1), ADK :Full, Cos2 800 nm 1e14 W/cm^2 2 cycles LP;
                      t0: n_t equidistant points over the pulse support +-N pi/omega, kd: n_kd
                      points in [-1, 1]; RK4 dt 0.1 to t_f = N pi/omega + 10; CTMC PMD 500 x 500
                      over +-1.5 a.u.  (n_t = 2000, n_kd = 50: 1e5 trajectories)
h_lp_scts_stratified  same atom and pulse, SCTS phase, n_t strata with one seeded jitter each and
                      n_k Gaussian k_perp (sigma 0.4 a.u.) per stratum, weights divided by the
                      sampling pdf; coherent PMD 1000 x 1000 over +-1.5 a.u.; mixed precision with
                      FP64 re-integration within 2 a.u. of the core
                      (n_t = 10000, n_k = 34400: 3.44e8 samples of which 1.0e8 electrons carry a
                      weight above the 1e-16 cutoff and are integrated; the rest sit at the pulse
                      edges and field zeros, where the ADK rate underflows the cutoff)
mol_lp_scts_wfat      single-centre molecule (Ip 0.45, Z 1, soft core 1.0) with a seeded 8 x 16 G00
                      table (seed 0xE7A10005), WFAT channel (0,0), Cos2 1300 nm 1e14 W/cm^2 4 cycles
                      LP, SCTS phase, stratified t0 x Gaussian k_perp (sigma 0.4 a.u.); coherent PMD
                      400 x 400 over +-2 a.u.; orientation (rot_α, rot_β, rot_γ)
ar_cp_spa_qtmc_3d     Ar (Tong-Lin SAE potential, Ip 0.579 a.u., Z 1, prefactor with l = 0), Cos2 800 nm
                      2e14 W/cm^2 2 cycles CP; SFA-SPA on a 3D (t0, kd, kz) grid n_t x n_k x n_k over the
                      pulse and [-1.5, 1.5]^2 a.u.; adaptive Tsit5 rtol 1e-8 (FP64; "mixed": FP32
                      stage derivatives, FP64 re-integration within 2 a.u.); QTMC; coherent 3D PMD
                      p_num^3 over +-1.5 a.u.  (n_t = 2000, n_k = 158: 5e7 samples, 256^3 bins)
bichromatic_scan_jobs counter-rotating bichromatic CP (Cos2 800 nm 8 cycles eps +1 and 400 nm 16 cycles
                      eps -1, equal intensities), H atom ADK (:Pre, :Jac) + SCTS, one job per intensity
                      of a log-spaced list, stratified t0 x Gaussian k_perp; 2D PMD 512 x 512 per job
wfat_orientation_jobs mol_lp_scts_wfat at n_rot orientations drawn uniformly on SO(3) from a seeded
                      Philox stream (seed 0xE7A10005)
"""
from __future__ import annotations

import math

from .lasers import Cos2Laser, AngFreq, CycNum
from .targets import TabulatedMolecule, get_atom, synthetic_g00_table


def _pulse():
    laser = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    half = CycNum(laser) * math.pi / AngFreq(laser)
    return laser, half


def h_lp_ctmc_grid(n_t: int = 2000, n_kd: int = 50, precision: str = "fp64") -> dict:
    laser, half = _pulse()
    return dict(init_cond_method="ADK", laser=laser, target=get_atom("H"), dimension=2,
                sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                final_p_max=(1.5, 1.5), final_p_num=(500, 500), ss_kd_max=1.0, ss_kd_num=n_kd,
                traj_phase_method="CTMC", rate_prefix="Full", sample_cutoff_limit=1e-16,
                traj_integrator="rk4", traj_dt=0.1, precision=precision)


def h_lp_scts_stratified(n_t: int = 10000, n_k: int = 34400, precision: str = "mixed", seed: int = 0xE7A10002,
                         p_num: int = 1000, restep_radius: float = 2.0) -> dict:
    """The 1e8-electron H-atom ADK+SCTS job: 10000 strata x 34400 samples, of which just over 1e8
    electrons carry a weight above the cutoff; FP64 re-integration of the electrons that come
    within 2 a.u. of the core."""
    laser, half = _pulse()
    return dict(init_cond_method="ADK", laser=laser, target=get_atom("H"), dimension=2,
                sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                final_p_max=(1.5, 1.5), final_p_num=(p_num, p_num), sample_stratified=True, mc_kt_num=n_k,
                kd_sigma=0.4, seed=seed, traj_phase_method="SCTS", rate_prefix="Full", sample_cutoff_limit=1e-16,
                traj_integrator="rk4", traj_dt=0.1, precision=precision, restep_radius=restep_radius)


def wfat_molecule(rot=(0.0, 0.0, 0.0), seed: int = 0xE7A10005) -> TabulatedMolecule:
    return TabulatedMolecule(Ip=0.45, charge=0, g00_coeff=synthetic_g00_table(seed), rot_α=rot[0], rot_β=rot[1],
                             rot_γ=rot[2], soft_core=1.0, name="tabulated single-centre molecule")


def mol_lp_scts_wfat(n_t: int = 4000, n_k: int = 1000, rot=(0.0, 0.0, 0.0), precision: str = "mixed",
                     seed: int = 0xE7A10005, p_num: int = 400, restep_radius: float = 0.0) -> dict:
    """The soft core (a = 1) removes the Coulomb singularity that the FP64 re-integration of close
    encounters exists for, so the mixed form re-integrates nothing by default (restep_radius 0)."""
    laser = Cos2Laser(peak_int=1e14, wave_len=1300.0, cyc_num=4, ellip=0.0)
    half = CycNum(laser) * math.pi / AngFreq(laser)
    return dict(init_cond_method="WFAT", laser=laser, target=wfat_molecule(rot, seed), dimension=2,
                sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                final_p_max=(2.0, 2.0), final_p_num=(p_num, p_num), sample_stratified=True, mc_kt_num=n_k,
                kd_sigma=0.4, seed=seed, traj_phase_method="SCTS", sample_cutoff_limit=1e-16,
                traj_integrator="rk4", traj_dt=0.1, precision=precision, restep_radius=restep_radius)


def ar_cp_spa_qtmc_3d(n_t: int = 2000, n_k: int = 158, rtol: float = 1e-8, p_num: int = 256,
                      method: str = "SPA", precision: str = "fp64") -> dict:
    """The SFA-SPA 3D grid job with adaptive Tsit5 at rtol 1e-8. precision "mixed" (FP32 stage
    derivatives, FP64 state and error control, FP64 re-integration near the core) is as close to
    a rtol 1e-10 FP64 run as FP64 at rtol 1e-8 is (PMD L1 3.5e-3 against 3.7e-3 at 256^3), but
    takes 18 % more steps and is no faster here, so FP64 is the default."""
    laser = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=1.0)
    half = CycNum(laser) * math.pi / AngFreq(laser)
    return dict(init_cond_method=method, laser=laser, target=get_atom("Ar", l=0), dimension=3,
                sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                final_p_max=(1.5, 1.5, 1.5), final_p_num=(p_num, p_num, p_num), ss_kd_max=1.5, ss_kd_num=n_k,
                ss_kz_max=1.5, ss_kz_num=n_k, traj_phase_method="QTMC", rate_prefix="Full",
                sample_cutoff_limit=1e-16, traj_integrator="tsit5", traj_rtol=rtol, precision=precision)


def bichromatic_laser(intensity: float) -> "BichromaticLaser":
    from .lasers import BichromaticLaser
    return BichromaticLaser(Cos2Laser(peak_int=intensity, wave_len=800.0, cyc_num=8, ellip=1.0),
                            Cos2Laser(peak_int=intensity, wave_len=400.0, cyc_num=16, ellip=-1.0), 0.0)


def bichromatic_scan_jobs(intensities, n_t: int = 2000, n_k=1000, precision: str = "mixed",
                          seed: int = 0xE7A10004, p_num: int = 512) -> list:
    """One keyword set per intensity; n_k may be a list (samples per stratum per job)."""
    half = 8 * math.pi / (45.563352525 / 800.0)
    n_ks = list(n_k) if hasattr(n_k, "__len__") else [n_k] * len(intensities)
    return [dict(init_cond_method="ADK", laser=bichromatic_laser(I), target=get_atom("H"), dimension=2,
                 sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                 final_p_max=(2.0, 2.0), final_p_num=(p_num, p_num), sample_stratified=True, mc_kt_num=int(nk),
                 kd_sigma=0.4, seed=seed + j, traj_phase_method="SCTS", rate_prefix={"Pre", "Jac"},
                 sample_cutoff_limit=1e-16, traj_integrator="rk4", traj_dt=0.1, precision=precision,
                 restep_radius=2.0)
            for j, (I, nk) in enumerate(zip(intensities, n_ks))]


def bichromatic_scan_counts(ion_prob, survivors, n_probe, total=1e8, lo=1e5, hi=2e7, n_t=2000):
    """Samples per stratum (n_k) of each intensity of the scan, so that job j integrates about
    clip(total * Y_j / sum Y, lo, hi) electrons: Y_j the job's ADK yield (ion_prob of a probe
    scan), and survivors_j / n_probe the fraction of its samples above the weight cutoff."""
    import numpy as np
    y = np.asarray(ion_prob, float)
    surv = np.maximum(np.asarray(survivors, float), 1.0) / n_probe
    want = np.clip(total * y / y.sum(), lo, hi)
    return [int(v) for v in np.maximum(1, np.ceil(want / (n_t * surv)))]


def so3_uniform(n: int, seed: int = 0xE7A10005):
    """n ZYZ Euler triples uniform on SO(3) (alpha, gamma uniform, cos beta uniform) from Philox4x64-10
    (numpy's Philox with key = seed)."""
    import numpy as np
    g = np.random.Generator(np.random.Philox(key=seed))
    u = g.random((n, 3))
    return [(2 * math.pi * a, math.acos(1 - 2 * b), 2 * math.pi * c) for a, b, c in u]


def wfat_orientation_jobs(n_rot: int = 64, n_t: int = 4000, n_k=1000, precision: str = "mixed",
                          p_num: int = 400) -> list:
    n_ks = list(n_k) if hasattr(n_k, "__len__") else [n_k] * n_rot
    return [mol_lp_scts_wfat(n_t=n_t, n_k=int(nk), rot=r, precision=precision, p_num=p_num)
            for r, nk in zip(so3_uniform(n_rot), n_ks)]
