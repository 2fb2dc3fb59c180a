"""Lasers: the reference's laser types and accessors (src/Lasers/).

Constructors mirror the keyword constructors of Cos2Laser.jl:44-74,
Cos4Laser.jl:44-74, GaussianLaser.jl:46-87 and BichromaticLaser.jl:32-35
(same argument names, defaults, assertions and warnings; numbers are taken in
the reference's numeric units: W/cm^2, nm, a.u., rad). `LaserFx(l)` etc.
return callables that evaluate the field ON THE DEVICE through
`etraj_laser_field`, i.e. the very device functions the integrator inlines.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import numpy as np

from . import _abi


class Laser:
    """Abstract laser (Lasers.jl:17)."""

    def _desc(self) -> _abi.LaserDesc:
        raise NotImplementedError


class MonochromaticLaser(Laser):
    """Abstract monochromatic laser (Lasers.jl:23)."""

    _kind = 0

    def _mono(self) -> _abi.MonoLaser:
        m = _abi.MonoLaser()
        m.kind = self._kind
        m.peak_int = self.peak_int
        m.wave_len = self.wave_len
        m.cyc_num = self._cyc
        m.ellip = self.ellip
        m.azi = self.azi
        m.cep = self.cep
        m.t_shift = self.t_shift
        return m

    def _desc(self) -> _abi.LaserDesc:
        d = _abi.LaserDesc()
        d.n_colors = 1
        d.color[0] = self._mono()
        return d


def _wave(wave_len, ang_freq, who):
    assert wave_len > 0 or ang_freq > 0, f"[{who}] Must specify either `wave_len` or `ang_freq`."
    if wave_len > 0 and ang_freq > 0:
        warnings.warn(f"[{who}] Both `wave_len` & `ang_freq` are specified, will use `wave_len`.")
    if wave_len == 0:
        wave_len = 45.563352525 / ang_freq
    return float(wave_len)


class _CosNLaser(MonochromaticLaser):
    _who = ""

    def __init__(self, *, peak_int, wave_len=0, ang_freq=0, cyc_num=0, duration=0, ellip, azi=0.0, cep=0.0,
                 t_shift=0.0):
        wave_len = _wave(wave_len, ang_freq, self._who)
        assert cyc_num > 0 or duration > 0, f"[{self._who}] Must specify either `cyc_num` or `duration`."
        if cyc_num > 0 and duration > 0:
            warnings.warn(f"[{self._who}] Both `cyc_num` & `duration` are specified, will use `cyc_num`.")
        if cyc_num == 0:
            cyc_num = duration / (2 * math.pi / (45.563352525 / wave_len))
        self.peak_int, self.wave_len, self.cyc_num = float(peak_int), wave_len, float(cyc_num)
        self.ellip, self.azi, self.cep, self.t_shift = float(ellip), float(azi), float(cep), float(t_shift)

    @property
    def _cyc(self):
        return self.cyc_num

    def __eq__(self, o):
        return type(o) is type(self) and self.__dict__ == o.__dict__

    def __repr__(self):
        return (f"[MonochromaticLaser] Envelope {self._env}, peak intensity {self.peak_int:.1e} W/cm², "
                f"wavelen={self.wave_len:g} nm, {self.cyc_num:g} cycle(s), ε={self.ellip:g}")


class Cos2Laser(_CosNLaser):
    """Cos2Laser.jl:7-74."""
    _kind = _abi.LASER_COS2
    _who = "Cos2Laser"
    _env = "cos²"


class Cos4Laser(_CosNLaser):
    """Cos4Laser.jl:7-74."""
    _kind = _abi.LASER_COS4
    _who = "Cos4Laser"
    _env = "cos⁴"


class GaussianLaser(MonochromaticLaser):
    """GaussianLaser.jl:7-87."""
    _kind = _abi.LASER_GAUSSIAN

    def __init__(self, *, peak_int, wave_len=0, ang_freq=0, spread_cyc_num=0, spread_duration=0, FWHM_duration=0,
                 ellip, azi=0.0, cep=0.0, t_shift=0.0):
        wave_len = _wave(wave_len, ang_freq, "GaussianLaser")
        assert spread_cyc_num > 0 or spread_duration > 0 or FWHM_duration > 0, \
            "[GaussianLaser] Must specify one in `spread_cyc_num`, `spread_duration` and `FWHM_duration`."
        if (spread_cyc_num > 0) + (spread_duration > 0) + (FWHM_duration > 0) > 1:
            warnings.warn("[GaussianLaser] More than one in `spread_cyc_num` & `spread_duration` & `FWHM_duration` "
                          "are specified, will use the first positive value.")
        T = 2 * math.pi / (45.563352525 / wave_len)
        if spread_cyc_num == 0:
            spread_cyc_num = (FWHM_duration / (2 * math.sqrt(math.log(2))) / T) if spread_duration == 0 \
                else spread_duration / T
        self.peak_int, self.wave_len, self.spread_cyc_num = float(peak_int), wave_len, float(spread_cyc_num)
        self.ellip, self.azi, self.cep, self.t_shift = float(ellip), float(azi), float(cep), float(t_shift)

    @property
    def _cyc(self):
        return self.spread_cyc_num

    def __eq__(self, o):
        return type(o) is type(self) and self.__dict__ == o.__dict__


@dataclass(eq=False)
class BichromaticLaser(Laser):
    """BichromaticLaser.jl:7-35: F(t) = F1(t) + F2(t - delay)."""
    l1: MonochromaticLaser
    l2: MonochromaticLaser
    delay: float = 0.0

    def __getitem__(self, i):
        if i == 1:
            return self.l1
        if i == 2:
            return self.l2
        raise IndexError("Index of the laser should be either 1 or 2.")

    def _desc(self) -> _abi.LaserDesc:
        d = _abi.LaserDesc()
        d.n_colors = 2
        d.delay21 = float(self.delay)
        d.color[0] = self.l1._mono()
        d.color[1] = self.l2._mono()
        return d


# ---- accessors (Lasers/exports.jl) ----------------------------------------------
def PeakInt(l): return l.peak_int
def WaveLen(l): return l.wave_len
def CycNum(l): return l.cyc_num
def SpreadCycNum(l): return l.spread_cyc_num
def Ellipticity(l): return l.ellip
def Azimuth(l): return l.azi
def CEP(l): return l.cep
def TimeShift(l): return l.t_shift
def AngFreq(l): return 45.563352525 / l.wave_len                       # Cos2Laser.jl:81
def Period(l): return 2 * math.pi / AngFreq(l)                          # :82
def LaserF0(l): return math.sqrt(l.peak_int / (1.0 + l.ellip ** 2) / 3.50944521e16)  # :85
def LaserA0(l): return LaserF0(l) / AngFreq(l)                          # :86
def KeldyshParameter(l, Ip): return AngFreq(l) * math.sqrt(2 * Ip) / LaserF0(l)     # :87
def SpreadDuration(l): return l.spread_cyc_num * Period(l)              # GaussianLaser.jl:120
def FWHM_Duration(l): return l.spread_cyc_num * Period(l) * (2 * math.sqrt(math.log(2)))
def Laser1(l): return l.l1
def Laser2(l): return l.l2
def Delay21(l): return l.delay


def _device_field(l: Laser, which: int):
    def f(t):
        ctx = _abi.default_context()
        ta = np.ascontiguousarray(np.atleast_1d(np.asarray(t, dtype=np.float64)))
        outs = [np.empty_like(ta) for _ in range(4)]
        ctx.check(ctx.lib.etraj_laser_field(ctx.handle, l._desc(), _abi.dptr(ta), ta.size,
                                            *[_abi.dptr(o) for o in outs]))
        r = outs[which]
        return r if np.ndim(t) else float(r[0])
    return f


def LaserFx(l: Laser):
    """t -> Fx(t), evaluated by the device field function (Cos2Laser.jl:125)."""
    return _device_field(l, 0)


def LaserFy(l: Laser):
    return _device_field(l, 1)


def LaserAx(l: Laser):
    return _device_field(l, 2)


def LaserAy(l: Laser):
    return _device_field(l, 3)


def Serialize(l: Laser) -> dict:
    """Cos2Laser.jl:192-204, BichromaticLaser.jl:101-109."""
    if isinstance(l, BichromaticLaser):
        return {"type": "BichromaticLaser", "laser1": Serialize(l.l1), "laser2": Serialize(l.l2),
                "delay21": l.delay}
    d = {"type": type(l).__name__}
    d.update(l.__dict__)
    return d
