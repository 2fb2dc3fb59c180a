"""CPU tests of the boundary and of the host-side mirror of the reference API
(no device calls)."""
import ctypes as C
import math
import os
import re
import warnings

import pytest

from etraj_hip import (_abi, Cos2Laser, Cos4Laser, GaussianLaser, BichromaticLaser, get_atom, HydrogenLikeAtom,
                       SAEAtom, init_sampler, LaserF0, LaserA0, AngFreq, Period, PeakInt, WaveLen, CycNum, Ellipticity,
                       Azimuth, CEP, TimeShift, SpreadDuration, FWHM_Duration, SpreadCycNum, Laser1, Laser2, Delay21,
                       IonPotential, AsympNuclCharge, AngularQuantumNumber, MagneticQuantumNumber, AsympCoeff,
                       SoftCore, TargetName, QuantizationAxisOrientaion, hartree_asymp_coeff)
from etraj_hip.samplers import resolve_rate_prefix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "etraj.h")).read()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(etraj_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == sorted(_abi.SIGNATURES)


def test_library_loads_and_exports_every_declared_symbol():
    lib = _abi.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.etraj_abi_version() == _abi.ABI_VERSION


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors agree with the C compiler's layout of include/etraj.h."""
    import subprocess
    structs = {"etraj_mono_laser": _abi.MonoLaser, "etraj_laser_desc": _abi.LaserDesc,
               "etraj_target_desc": _abi.TargetDesc, "etraj_sampler_desc": _abi.SamplerDesc,
               "etraj_run_desc": _abi.RunDesc, "etraj_stats": _abi.Stats, "etraj_tuning": _abi.Tuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "etraj.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_batch_max_size_without_device():
    lib = _abi.load_library()
    d = _abi.SamplerDesc(dimension=3, sampling=_abi.SAMPLING_STEP, ss_kd_num=30, ss_kz_num=40)
    assert lib.etraj_batch_max_size(d) == 1200
    d = _abi.SamplerDesc(dimension=2, sampling=_abi.SAMPLING_STRATIFIED, mc_kt_num=10000)
    assert lib.etraj_batch_max_size(d) == 10000


# ---- Lasers_test.jl ------------------------------------------------------------
def test_cos2_constructor_equivalences_and_properties():
    l1 = Cos2Laser(peak_int=4e14, wave_len=800.0, cyc_num=2.0, ellip=1.0, azi=math.pi / 2, cep=math.pi, t_shift=10.0)
    l2 = Cos2Laser(peak_int=4e14, ang_freq=0.05695419065625, cyc_num=2.0, ellip=1.0, azi=math.pi / 2, cep=math.pi,
                   t_shift=10.0)
    l3 = Cos2Laser(peak_int=4e14, wave_len=800.0, duration=220.63996467273427, ellip=1.0, azi=math.pi / 2,
                   cep=math.pi, t_shift=10.0)
    assert l1 == l2 == l3                                          # Lasers_test.jl:19-21
    assert PeakInt(l1) == 4e14 and WaveLen(l1) == 800.0 and CycNum(l1) == 2.0
    assert Ellipticity(l1) == 1.0 and Azimuth(l1) == math.pi / 2 and CEP(l1) == math.pi and TimeShift(l1) == 10.0
    assert AngFreq(l1) == 45.563352525 / WaveLen(l1)
    assert Period(l1) == 2 * math.pi / AngFreq(l1)
    assert LaserF0(l1) == math.sqrt(PeakInt(l1) / (1.0 + Ellipticity(l1) ** 2) / 3.50944521e16)
    assert LaserA0(l1) == LaserF0(l1) / AngFreq(l1)


def test_gaussian_constructor_equivalences():
    kw = dict(peak_int=4e14, wave_len=800.0, ellip=1.0, azi=math.pi / 2, cep=math.pi, t_shift=10.0)
    l1 = GaussianLaser(spread_cyc_num=2.0, **kw)
    l3 = GaussianLaser(spread_duration=220.63996467273427, **kw)
    l4 = GaussianLaser(FWHM_duration=367.38963998791286, **kw)
    assert SpreadCycNum(l1) == 2.0
    assert SpreadCycNum(l3) == pytest.approx(2.0, rel=1e-15)
    assert SpreadCycNum(l4) == pytest.approx(2.0, rel=1e-15)
    assert SpreadDuration(l1) == SpreadCycNum(l1) * Period(l1)
    assert FWHM_Duration(l1) == SpreadDuration(l1) * (2 * math.sqrt(math.log(2)))


def test_bichromatic_accessors():
    a = Cos4Laser(peak_int=4e14, wave_len=800.0, cyc_num=6.0, ellip=1.0)
    b = Cos4Laser(peak_int=4e14, wave_len=400.0, cyc_num=6.0, ellip=-1.0)
    l = BichromaticLaser(a, b, 50.0)
    assert Laser1(l) is l[1] is a and Laser2(l) is l[2] is b and Delay21(l) == 50.0
    d = l._desc()
    assert d.n_colors == 2 and d.delay21 == 50.0 and d.color[1].wave_len == 400.0


def test_laser_constructor_assertions():
    with pytest.raises(AssertionError):
        Cos2Laser(peak_int=1e14, cyc_num=2, ellip=0.0)
    with pytest.raises(AssertionError):
        Cos2Laser(peak_int=1e14, wave_len=800.0, ellip=0.0)


# ---- Targets_test.jl -----------------------------------------------------------
def test_hydrogen_like_atom():
    t1 = HydrogenLikeAtom(Ip=0.5, Z=1, l=0, m=0, asymp_coeff="hartree", quan_ax_θ=math.pi / 2, quan_ax_ϕ=math.pi,
                          soft_core=1e-10, name="H")
    t3 = get_atom("H", quan_ax_θ=math.pi / 2, quan_ax_ϕ=math.pi)
    assert t1 == t3
    assert IonPotential(t1) == 0.5 and AsympNuclCharge(t1) == 1 and AngularQuantumNumber(t1) == 0
    assert MagneticQuantumNumber(t1) == 0 and QuantizationAxisOrientaion(t1) == (math.pi / 2, math.pi)
    assert AsympCoeff(t1) == 1.0 and SoftCore(t1) == 1e-10 and TargetName(t1) == "H"


def test_sae_atom_hartree_coefficient():
    t1 = get_atom("He")
    assert AsympCoeff(t1) == pytest.approx(0.9124458182520485, rel=1e-15)   # Targets_test.jl:43,56
    assert t1.a1 == 1.230723 and t1.b3 == 0.4804286
    with pytest.raises(AssertionError):
        SAEAtom(Ip=0.5, Z=1, soft_core=0.0)


def test_atom_database_lookups():
    assert get_atom("Ar").Ip == 0.579155055 and get_atom("Ar").l == 1
    assert get_atom("Xe", 1).name == "Xe⁺" and get_atom("Xe", 1).nucl_charge == 2
    with pytest.raises(KeyError):
        get_atom("Xe", 2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert hartree_asymp_coeff(1, 0.5, 0) == 1.0


# ---- ElectronSamplers ----------------------------------------------------------
def test_rate_prefix_resolution():
    l = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    assert resolve_rate_prefix("Full", l) == _abi.PREFIX_PRECC | _abi.PREFIX_JAC
    assert resolve_rate_prefix(":Exp", l) == 0
    assert resolve_rate_prefix({"Pre", "Jac"}, l) == _abi.PREFIX_PRE | _abi.PREFIX_JAC
    with pytest.raises(ValueError):
        resolve_rate_prefix({"Pre", "PreCC"}, l)
    with pytest.raises(ValueError):
        resolve_rate_prefix("Foo", l)
    b = BichromaticLaser(l, l, 0.0)
    with pytest.warns(UserWarning):
        assert resolve_rate_prefix("Full", b) == _abi.PREFIX_PRE | _abi.PREFIX_JAC   # ADKSampler.jl:72-77


def test_init_sampler_dispatch():
    l = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    kw = dict(laser=l, target=get_atom("H"), dimension=2, sample_t_intv=(-100, 100), sample_t_num=10,
              ss_kd_max=1.0, ss_kd_num=10)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method=":ADK", **kw)
    assert sp.desc.sampling == _abi.SAMPLING_STEP and sp.rows == 6
    for m, cls in (("SPA", "SPASampler"), ("MOSFA", "SPASampler"), ("SPANE", "SPANESampler"),
                   ("SFAAE", "SPANESampler")):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sp = init_sampler(init_cond_method=m, **kw)
        assert type(sp).__name__ == cls and sp.desc.init_method in (_abi.INIT_SPA, _abi.INIT_SPANE)
    with pytest.raises(TypeError):              # WFAT only supports molecular targets
        init_sampler(init_cond_method="WFAT", **kw)
    with pytest.raises(ValueError):
        init_sampler(init_cond_method="XYZ", **kw)
    with pytest.raises(AssertionError):
        init_sampler(init_cond_method="ADK", **dict(kw, ss_kd_num=0))


def test_no_device_means_loud_failure():
    """Without a HIP device the engine refuses; it never falls back to the CPU."""
    try:
        n = C.c_int()
        hip = C.CDLL("libamdhip64.so")
        has_dev = hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        has_dev = False
    if has_dev:
        pytest.skip("a HIP device is present")
    with pytest.raises(_abi.EngineUnavailable):
        _abi.Context(0)
