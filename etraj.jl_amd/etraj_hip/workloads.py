"""Named workloads of the engine (keyword sets for perform_traj_simulation).

h_lp_ctmc_grid        H atom (Ip 0.5, Z 1), ADK :Full, Cos2 800 nm 1e14 W/cm^2 2 cycles LP;
                      t0: n_t equidistant points over the pulse support +-N pi/omega, kd: n_kd
                      points in [-1, 1]; RK4 dt 0.1 to t_f = N pi/omega + 10; CTMC PMD 500 x 500
                      over +-1.5 a.u.  (n_t = 2000, n_kd = 50: 1e5 trajectories)
h_lp_scts_stratified  same atom and pulse, SCTS phase, n_t strata with one seeded jitter each and
                      n_k Gaussian k_perp (sigma 0.4 a.u.) per stratum, weights divided by the
                      sampling pdf; coherent PMD 1000 x 1000 over +-1.5 a.u.
                      (n_t = 10000, n_k = 32000: 3.2e8 samples of which ~1.0e8 electrons carry a
                      weight above the 1e-16 cutoff and are integrated; the rest sit at the pulse
                      edges and field zeros, where the ADK rate underflows the cutoff)
mol_lp_scts_wfat      single-centre molecule (Ip 0.45, Z 1, soft core 1.0) with a seeded 8 x 16 G00
                      table (seed 0xE7A10005), WFAT channel (0,0), Cos2 1300 nm 1e14 W/cm^2 4 cycles
                      LP, SCTS phase, stratified t0 x Gaussian k_perp (sigma 0.4 a.u.); coherent PMD
                      400 x 400 over +-2 a.u.; orientation (rot_α, rot_β, rot_γ)
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


def h_lp_scts_stratified(n_t: int = 10000, n_k: int = 32000, precision: str = "mixed", seed: int = 0xE7A10002,
                         p_num: int = 1000) -> dict:
    laser, half = _pulse()
    return dict(init_cond_method="ADK", laser=laser, target=get_atom("H"), dimension=2,
                sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                final_p_max=(1.5, 1.5), final_p_num=(p_num, p_num), sample_stratified=True, mc_kt_num=n_k,
                kd_sigma=0.4, seed=seed, traj_phase_method="SCTS", rate_prefix="Full", sample_cutoff_limit=1e-16,
                traj_integrator="rk4", traj_dt=0.1, precision=precision, restep_radius=2.0)


def wfat_molecule(rot=(0.0, 0.0, 0.0), seed: int = 0xE7A10005) -> TabulatedMolecule:
    return TabulatedMolecule(Ip=0.45, charge=0, g00_coeff=synthetic_g00_table(seed), rot_α=rot[0], rot_β=rot[1],
                             rot_γ=rot[2], soft_core=1.0, name="tabulated single-centre molecule")


def mol_lp_scts_wfat(n_t: int = 4000, n_k: int = 1000, rot=(0.0, 0.0, 0.0), precision: str = "mixed",
                     seed: int = 0xE7A10005, p_num: int = 400) -> dict:
    laser = Cos2Laser(peak_int=1e14, wave_len=1300.0, cyc_num=4, ellip=0.0)
    half = CycNum(laser) * math.pi / AngFreq(laser)
    return dict(init_cond_method="WFAT", laser=laser, target=wfat_molecule(rot, seed), dimension=2,
                sample_t_intv=(-half, half), sample_t_num=n_t, traj_t_final=half + 10.0,
                final_p_max=(2.0, 2.0), final_p_num=(p_num, p_num), sample_stratified=True, mc_kt_num=n_k,
                kd_sigma=0.4, seed=seed, traj_phase_method="SCTS", sample_cutoff_limit=1e-16,
                traj_integrator="rk4", traj_dt=0.1, precision=precision, restep_radius=2.0)
