"""ElectronSamplers: init_sampler / batch_num / gen_electron_batch
(src/ElectronSamplers/ElectronSamplers.jl:9-26, ADKSampler.jl, SPASampler.jl,
SPANESampler.jl, WFATSampler.jl).

`gen_electron_batch` runs the device sampler kernel and returns the same
rows x N column-major matrix the reference returns (rows x,y,kx,ky,t0,rate
[,phase] in 2D; x,y,z,kx,ky,kz,t0,rate[,phase] in 3D), or None for an empty
batch. Monte-Carlo sampling is seeded (Philox4x64-10 keyed by `seed`)
instead of the reference's unseeded global rand().
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _abi
from .lasers import BichromaticLaser, MonochromaticLaser, KeldyshParameter
from .targets import SAEAtomBase, IonPotential

_SPA = {"SPA", "MOSPA", "SFA", "MOSFA"}
_SPANE = {"SPANE", "MOSPANE", "SFAAE", "MOSFAAE"}
_ADK = {"ADK", "MOADK"}


def _sym(x):
    return x.lstrip(":") if isinstance(x, str) else x


def resolve_rate_prefix(rate_prefix, laser, who="ADKSampler") -> int:
    """ADKSampler.jl:49-77 (and SPASampler.jl:44-71, SPANESampler.jl) -> ETRAJ_PREFIX_* bits."""
    names = {"Pre": _abi.PREFIX_PRE, "PreCC": _abi.PREFIX_PRECC, "Jac": _abi.PREFIX_JAC}
    if isinstance(rate_prefix, str):
        p = _sym(rate_prefix)
        if p == "Exp":
            s = set()
        elif p == "Full":
            s = {"PreCC", "Jac"}
        elif p in names:
            s = {p}
        else:
            raise ValueError(f"[{who}] Undefined tunneling rate prefix [{rate_prefix}].")
    else:
        s = {_sym(p) for p in rate_prefix}
        if not all(p in names for p in s):
            raise ValueError(f"[{who}] Undefined tunneling rate prefix [{rate_prefix}].")
        if "Pre" in s and "PreCC" in s:
            raise ValueError(f"[{who}] Rate prefixes [Pre] & [PreCC] conflict.")
    if "PreCC" in s and not isinstance(laser, MonochromaticLaser):
        warnings.warn(f"[{who}] Laser is not monochromatic, Coulomb correction in rate prefix is unavailable.")
        s = (s - {"PreCC"}) | {"Pre"}
    bits = 0
    for p in s:
        bits |= names[p]
    return bits


def _fill_sampling(d, who, dimension, sample_t_intv, sample_t_num, sample_cutoff_limit, sample_monte_carlo,
                   sample_stratified, ss_kd_max, ss_kd_num, ss_kz_max, ss_kz_num, mc_kt_num, mc_kd_max, mc_kz_max,
                   kd_sigma, kz_sigma, seed):
    """Sampling-mode fields and their assertions, shared by the samplers (ADKSampler.jl:78-110,
    WFATSampler.jl:45-72)."""
    assert sample_t_num > 0, f"[{who}] Invalid time sample number {sample_t_num}."
    assert sample_cutoff_limit >= 0, f"[{who}] Invalid cut-off limit {sample_cutoff_limit}."
    d.sample_t_num = int(sample_t_num)
    d.sample_t_lo, d.sample_t_hi = float(sample_t_intv[0]), float(sample_t_intv[1])
    d.cutoff_limit = float(sample_cutoff_limit)
    d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    if sample_stratified:
        d.sampling = _abi.SAMPLING_STRATIFIED
        d.mc_kt_num = int(mc_kt_num)
        d.kd_sigma, d.kz_sigma = float(kd_sigma), float(kz_sigma)
        assert mc_kt_num > 0 and kd_sigma > 0, "stratified sampling needs mc_kt_num > 0 and kd_sigma > 0"
    elif sample_monte_carlo:
        d.sampling = _abi.SAMPLING_MONTECARLO
        assert sample_t_intv[0] < sample_t_intv[1], f"[{who}] Invalid sampling time interval {sample_t_intv}."
        assert mc_kt_num > 0, f"[{who}] Invalid sampling kt_num {mc_kt_num}."
        if dimension == 3:
            assert mc_kd_max > 0 or mc_kz_max > 0, f"[{who}] Invalid kd,kz sample boundaries."
        else:
            assert mc_kd_max > 0, f"[{who}] Invalid kd sample boundaries {mc_kd_max}."
        d.mc_kt_num, d.mc_kd_max, d.mc_kz_max = int(mc_kt_num), float(mc_kd_max), float(mc_kz_max)
    else:
        d.sampling = _abi.SAMPLING_STEP
        if dimension == 3:
            assert ss_kd_num > 0 and ss_kz_num > 0, f"[{who}] Invalid kd,kz sample number {ss_kd_num},{ss_kz_num}."
            assert ss_kd_max > 0 or ss_kz_max > 0, f"[{who}] Invalid kd,kz sample boundaries."
            d.ss_kz_num, d.ss_kz_max = int(ss_kz_num), float(ss_kz_max)
        else:
            assert ss_kd_num > 0, f"[{who}] Invalid kd sample number {ss_kd_num}."
            assert ss_kd_max > 0, f"[{who}] Invalid kd sample boundaries {ss_kd_max}."
            d.ss_kz_num, d.ss_kz_max = 1, 0.0
        d.ss_kd_num, d.ss_kd_max = int(ss_kd_num), float(ss_kd_max)


def _keldysh_warning(who, laser, Ip):
    """ADKSampler.jl:116-131, WFATSampler.jl:80-95 (the second branch of each is unreachable there)."""
    if isinstance(laser, MonochromaticLaser):
        g = KeldyshParameter(laser, Ip)
    elif isinstance(laser, BichromaticLaser):
        g = max(KeldyshParameter(laser.l1, Ip), KeldyshParameter(laser.l2, Ip))
    else:
        return
    if g >= 0.6:
        warnings.warn(f"[{who}] Keldysh parameter γ={g:.4f}, adiabatic (tunneling) condition [γ<<1] "
                      "not sufficiently satisfied.")


class ElectronSampler:
    """ElectronSamplers.jl:28."""


class ADKSampler(ElectronSampler):
    """ADKSampler.jl:3-150 (atoms and GenericMolecule); device sampling."""
    _method = _abi.INIT_ADK
    _who = "ADKSampler"

    def __init__(self, *, laser, target, dimension, sample_t_intv, sample_t_num, sample_cutoff_limit=1e-16,
                 sample_monte_carlo=False, traj_phase_method="CTMC", rate_prefix="Full", ss_kd_max=0.0, ss_kd_num=0, ss_kz_max=0.0,
                 ss_kz_num=0, mc_kt_num=0, mc_kd_max=0.0, mc_kz_max=0.0, mol_orbit_ridx=0, seed=0,
                 sample_stratified=False, kd_sigma=0.0, kz_sigma=0.0, **_):
        from .targets import GenericMolecule
        who = self._who
        if not isinstance(target, (SAEAtomBase, GenericMolecule)):
            raise NotImplementedError(f"[{who}] the engine samples atoms (SAEAtomBase) and GenericMolecule targets")
        pm = _sym(traj_phase_method)
        if pm not in _abi.PHASE_METHODS:
            raise ValueError(f"[{who}] Undefined phase method [{traj_phase_method}].")
        self.laser, self.target = laser, target
        d = _abi.SamplerDesc()
        d.init_method = self._method
        d.dimension = int(dimension)
        d.phase_method = _abi.PHASE_METHODS[pm]
        d.rate_prefix = resolve_rate_prefix(rate_prefix, laser, who)
        _fill_sampling(d, who, dimension, sample_t_intv, sample_t_num, sample_cutoff_limit,
                       sample_monte_carlo, sample_stratified, ss_kd_max, ss_kd_num, ss_kz_max, ss_kz_num, mc_kt_num,
                       mc_kd_max, mc_kz_max, kd_sigma, kz_sigma, seed)
        Ip = IonPotential(target)
        if isinstance(target, GenericMolecule):     # MO-ADK / MO-SFA: the orbital's C_lm (ADKSampler.jl:188-265)
            Ip = target.ion_potential(mol_orbit_ridx)
            if Ip <= 0:
                raise ValueError(f"[{who}] The energy of the ionizing orbit is non-negative.")
            C = target.MolAsympCoeff(mol_orbit_ridx)
            flat = C.ravel(order="F")               # Julia's column-major memory of C[l, m + l]
            coeff = np.ascontiguousarray(np.column_stack([flat.real, flat.imag]).ravel())
            d.asymp_lmax = C.shape[0] - 1
            d.asymp_coeff = _abi.dptr(coeff)
            d._keep_asymp = coeff                   # the descriptor points into it
            d.orbit_ip = Ip
            for i, a in enumerate(target.rot):
                d.mol_rot[i] = a
        if self._method == _abi.INIT_ADK:
            _keldysh_warning(who, laser, Ip)
        self.desc = d

    @property
    def rows(self) -> int:
        d = self.desc
        return (8 if d.phase_method == _abi.PHASE_CTMC else 9) if d.dimension == 3 else \
            (6 if d.phase_method == _abi.PHASE_CTMC else 7)

    @property
    def t_samples(self) -> np.ndarray:
        ctx = _abi.default_context()
        t = np.empty(self.desc.sample_t_num)
        ctx.check(ctx.lib.etraj_batch_times(ctx.handle, self.desc, _abi.dptr(t)))
        return t


class SPASampler(ADKSampler):
    """SPASampler.jl:1-148 (SFA-SPA; atoms with l = 0); device sampling. Each
    sample solves the complex saddle-point equation for its tunnelling time."""
    _method = _abi.INIT_SPA
    _who = "SPASampler"


class SPANESampler(ADKSampler):
    """SPANESampler.jl:1-148 (SFA-SPA with non-adiabatic expansion; atoms with
    l = 0); device sampling."""
    _method = _abi.INIT_SPANE
    _who = "SPANESampler"


class WFATSampler(ADKSampler):
    """WFATSampler.jl:3-117 (molecules, channel (0,0)); device sampling.

    The structure factor comes from the target's G00 table
    (`TabulatedMolecule`). The reference accepts the :CTMC phase method only
    (WFATSampler.jl:41-44); :QTMC and :SCTS are accepted here as an extension
    (the initial phase row is 0, as in ADK)."""

    def __init__(self, *, laser, target, dimension, sample_t_intv, sample_t_num, sample_cutoff_limit=1e-16,
                 sample_monte_carlo=False, traj_phase_method="CTMC", mol_orbit_ridx=0, ss_kd_max=0.0, ss_kd_num=0,
                 ss_kz_max=0.0, ss_kz_num=0, mc_kt_num=0, mc_kd_max=0.0, mc_kz_max=0.0, seed=0,
                 sample_stratified=False, kd_sigma=0.0, kz_sigma=0.0, **_):
        from .targets import MoleculeBase, MolRotation
        if not isinstance(target, MoleculeBase):
            raise TypeError("[WFATSampler] WFAT only supports molecular targets (MoleculeBase).")
        pm = _sym(traj_phase_method)
        if pm not in _abi.PHASE_METHODS:
            raise ValueError(f"[WFATSampler] Unsupported phase method [{traj_phase_method}].")
        from .targets import GenericMolecule
        ip_orbit = target.ion_potential(mol_orbit_ridx) if isinstance(target, GenericMolecule) else IonPotential(target)
        if ip_orbit <= 0:                                        # WFATSampler.jl:77-79
            raise ValueError("[WFATSampler] The energy of the ionizing orbit is non-negative.")
        self.laser, self.target = laser, target
        d = _abi.SamplerDesc()
        d.init_method = _abi.INIT_WFAT
        d.dimension = int(dimension)
        d.phase_method = _abi.PHASE_METHODS[pm]
        _fill_sampling(d, "WFATSampler", dimension, sample_t_intv, sample_t_num, sample_cutoff_limit,
                       sample_monte_carlo, sample_stratified, ss_kd_max, ss_kd_num, ss_kz_max, ss_kz_num, mc_kt_num,
                       mc_kd_max, mc_kz_max, kd_sigma, kz_sigma, seed)
        for i, a in enumerate(MolRotation(target)):
            d.mol_rot[i] = a
        if isinstance(target, GenericMolecule):     # every channel of the WFAT integrals
            mu, intdata = target.wfat_data(mol_orbit_ridx)
            # Julia's column-major memory of intdata[nxi, m, l, m'] as (re, im) pairs
            flat = intdata.ravel(order="F")
            coeff = np.ascontiguousarray(np.column_stack([flat.real, flat.imag]).ravel())
            d.wfat_form = _abi.WFAT_INTDATA
            for k in range(4):
                d.wfat_dims[k] = intdata.shape[k]
            for k in range(3):
                d.wfat_mu[k] = mu[k]
            d.orbit_ip = ip_orbit
            d.wfat_na, d.wfat_nb = intdata.shape[2], intdata.shape[3]
        else:                                       # channel (0, 0) from a G00 table
            coeff = target.coeff_table()
            d.wfat_form = _abi.WFAT_TABLE
            d.wfat_na, d.wfat_nb = target.g00_coeff.shape
        d.wfat_coeff = _abi.dptr(coeff)
        d._keep_coeff = coeff                       # the descriptor points into it
        _keldysh_warning("WFATSampler", laser, IonPotential(target))
        self.desc = d


def init_sampler(**kwargs) -> ElectronSampler:
    """ElectronSamplers.jl:11-24."""
    m = _sym(kwargs["init_cond_method"])
    if m in _ADK:
        return ADKSampler(**kwargs)
    if m == "WFAT":
        return WFATSampler(**kwargs)
    if m in _SPA:
        return SPASampler(**kwargs)
    if m in _SPANE:
        return SPANESampler(**kwargs)
    raise ValueError(f"[ElectronSamplers] Undefined initial condition method [{m}].")


def batch_num(sp: ElectronSampler) -> int:
    """ADKSampler.jl:153-155."""
    return int(sp.desc.sample_t_num)


def batch_max_size(sp: ElectronSampler) -> int:
    """ADKSampler.jl:158-164."""
    return int(_abi.load_library().etraj_batch_max_size(sp.desc))


def gen_electron_batch(sp: ElectronSampler, batch_id: int):
    """ADKSampler.jl:167-396 on the device; batch_id is 1-based as in Julia."""
    ctx = _abi.default_context()
    cap = max(batch_max_size(sp), 1)
    out = np.empty((cap, sp.rows))  # C-order (cap, rows) == Julia column-major rows x cap
    cnt = _abi.C.c_int64()
    ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, sp.laser._desc(), sp.target._desc(), sp.desc,
                                               int(batch_id), _abi.dptr(out), cap, _abi.C.byref(cnt)))
    n = cnt.value
    if n == 0:
        return None
    return np.asfortranarray(out[:n].T)
