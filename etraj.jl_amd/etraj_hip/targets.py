"""Targets: hydrogen-like and SAE atoms (src/Targets/).

Constructors mirror HydrogenLikeAtom.jl:49-64 and SAEAtom.jl:56-72 (same
keywords, defaults and assertions; Ip in a.u.). `TargetPotential(t)` and
`TargetForce(t)` evaluate the device functions through `etraj_target_eval`.
`get_atom` uses the Tong-Lin SAE parameter table of AtomDatabase.jl:54-275
(J. Phys. B 38, 2593 (2005)).
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from . import _abi


class Target:
    """Target.jl: abstract target."""


class SAEAtomBase(Target):
    def _desc(self) -> _abi.TargetDesc:
        d = _abi.TargetDesc()
        d.kind = self._kind
        d.nucl_charge = self.nucl_charge
        d.l, d.m = self.l, self.m
        d.Ip = self.Ip
        d.asymp_coeff = self.asymp_coeff
        d.quan_ax_theta, d.quan_ax_phi = self.quan_ax_θ, self.quan_ax_ϕ
        d.soft_core = self.soft_core
        for k in ("a1", "b1", "a2", "b2", "a3", "b3"):
            setattr(d, k, getattr(self, k, 0.0))
        return d

    def __eq__(self, o):
        return type(o) is type(self) and self.__dict__ == o.__dict__


def hartree_asymp_coeff(Z, Ip, l):
    """SAEAtomBase_shared.jl:3-10."""
    if Z / math.sqrt(2 * Ip) <= l:
        warnings.warn("[hartree_asymp_coeff] Z/κ<=l, unable to determine the asymptotic coefficient using Hartree "
                      "formula. Using the l=0 case instead.")
        l = 0
    n = Z / math.sqrt(2 * Ip)
    return 2 ** (n - 1) / math.sqrt(n * math.gamma(n + l + 1) * math.gamma(n - l))


def _coeff(asymp_coeff, Z, Ip, l, who):
    assert asymp_coeff == "hartree" or (not isinstance(asymp_coeff, str) and asymp_coeff > 0), \
        f"[{who}] asymp_coeff should be either `:hartree` or a positive number."
    return hartree_asymp_coeff(Z, Ip, l) if asymp_coeff == "hartree" else float(asymp_coeff)


class HydrogenLikeAtom(SAEAtomBase):
    """HydrogenLikeAtom.jl:7-64: V = -Z (r^2 + soft_core)^(-1/2)."""
    _kind = _abi.TARGET_HYDROGEN_LIKE

    def __init__(self, *, Ip, Z, l=0, m=0, asymp_coeff="hartree", quan_ax_θ=0.0, quan_ax_ϕ=0.0, soft_core=1e-10,
                 name="[NA]"):
        assert Ip > 0, "[HydrogenLikeAtom] `Ip` should be positive."
        assert l >= 0 and l >= abs(m), "[HydrogenLikeAtom] Invalid (l,m)."
        assert soft_core >= 0, "[HydrogenLikeAtom] `soft_core` should be non-negative."
        self.Ip, self.nucl_charge, self.l, self.m = float(Ip), int(Z), int(l), int(m)
        self.asymp_coeff = _coeff(asymp_coeff, Z, Ip, l, "HydrogenLikeAtom")
        self.quan_ax_θ, self.quan_ax_ϕ, self.soft_core, self.name = float(quan_ax_θ), float(quan_ax_ϕ), \
            float(soft_core), name

    def __repr__(self):
        return f"[HydrogenLikeAtom] Atom {self.name}, Ip={self.Ip:.4f}, Z={self.nucl_charge}"


class SAEAtom(SAEAtomBase):
    """SAEAtom.jl:7-72: Tong-Lin model potential."""
    _kind = _abi.TARGET_SAE

    def __init__(self, *, Ip, Z, l=0, m=0, asymp_coeff="hartree", quan_ax_θ=0.0, quan_ax_ϕ=0.0, a1=0.0, b1=0.0,
                 a2=0.0, b2=0.0, a3=0.0, b3=0.0, soft_core=1e-10, name="[NA]"):
        assert Ip > 0, "[SAEAtom] Ip should be positive."
        assert l >= 0 and l >= abs(m), "[SAEAtom] Invalid (l,m)."
        assert b1 >= 0 and b2 >= 0 and b3 >= 0, "[SAEAtom] b1,b2,b3 should be non-negative."
        assert soft_core > 0, "[SAEAtom] soft_core should be positive."
        self.Ip, self.nucl_charge, self.l, self.m = float(Ip), int(Z), int(l), int(m)
        self.asymp_coeff = _coeff(asymp_coeff, Z, Ip, l, "SAEAtom")
        self.quan_ax_θ, self.quan_ax_ϕ = float(quan_ax_θ), float(quan_ax_ϕ)
        self.a1, self.b1, self.a2, self.b2, self.a3, self.b3 = map(float, (a1, b1, a2, b2, a3, b3))
        self.soft_core, self.name = float(soft_core), name

    def __repr__(self):
        return f"[SAEAtom] Atom {self.name}, Ip={self.Ip:.4f}, Z={self.nucl_charge}"


class MoleculeBase(Target):
    """Targets.jl: abstract molecular target."""


class TabulatedMolecule(MoleculeBase):
    """A molecule reduced to what the WFAT trajectory path reads from
    GenericMolecule (GenericMolecule.jl): the ionizing orbital's Ip, the
    asymptotic charge Z = charge + 1, the orientation (rot_α, rot_β, rot_γ)
    (z-y'-z'' Euler angles in radians) and the channel-(0,0) structure factor
    G00(β, γ) = Σ_{a,b} c_ab P_a(cos β) e^{-i b γ}, given as the complex
    table `g00_coeff[a, b]` in place of the WFAT integrals of a
    quantum-chemistry run (GenericMolecule.jl:236-260).

    The EOM is GenericMolecule's single-centre soft-core Coulomb potential
    V = -Z (r^2 + 1)^(-1/2) (GenericMolecule.jl:453-454); `soft_core` may be
    changed.
    """

    def __init__(self, *, Ip, charge=0, g00_coeff, rot_α=0.0, rot_β=0.0, rot_γ=0.0, soft_core=1.0, name="[NA]"):
        assert Ip > 0, "[TabulatedMolecule] the energy of the ionizing orbit is non-negative."
        assert soft_core >= 0, "[TabulatedMolecule] `soft_core` should be non-negative."
        c = np.asarray(g00_coeff, dtype=np.complex128)
        assert c.ndim == 2 and c.size > 0, "[TabulatedMolecule] g00_coeff should be a non-empty 2D table."
        self.Ip, self.charge, self.nucl_charge = float(Ip), int(charge), int(charge) + 1
        self.g00_coeff = c
        self.rot = (float(rot_α), float(rot_β), float(rot_γ))
        self.soft_core, self.name = float(soft_core), name
        self.l = self.m = 0
        self.asymp_coeff = 1.0
        self.quan_ax_θ = self.quan_ax_ϕ = 0.0

    def _desc(self) -> _abi.TargetDesc:
        d = _abi.TargetDesc()
        d.kind = _abi.TARGET_HYDROGEN_LIKE
        d.nucl_charge = self.nucl_charge
        d.Ip = self.Ip
        d.asymp_coeff = 1.0
        d.soft_core = self.soft_core
        return d

    def coeff_table(self) -> np.ndarray:
        """(re, im) pairs, index a * nb + b (the sampler descriptor's layout)."""
        return np.ascontiguousarray(np.stack([self.g00_coeff.real, self.g00_coeff.imag], axis=-1).ravel())

    def with_rotation(self, rot_α, rot_β, rot_γ) -> "TabulatedMolecule":
        return TabulatedMolecule(Ip=self.Ip, charge=self.charge, g00_coeff=self.g00_coeff, rot_α=rot_α, rot_β=rot_β,
                                 rot_γ=rot_γ, soft_core=self.soft_core, name=self.name)

    def __eq__(self, o):
        return (type(o) is type(self) and self.Ip == o.Ip and self.charge == o.charge and self.rot == o.rot
                and self.soft_core == o.soft_core and np.array_equal(self.g00_coeff, o.g00_coeff))

    def __repr__(self):
        return f"[TabulatedMolecule] {self.name}, Ip={self.Ip:.4f}, Z={self.nucl_charge}, rot={self.rot}"


class GenericMolecule(MoleculeBase):
    """GenericMolecule.jl's molecule as far as the trajectory path reads it:
    the data of `LoadMolecule` (GenericMolecule.jl:99-119), i.e. a molecule
    file of eTraj.jl's database (`get_mol`) or the same fields as a dict.

    * IonPotential = -max E(HOMO) over the spins (:438-443), AsympNuclCharge
      = charge + 1 (:450); the EOM is the single-centre soft-core Coulomb
      potential -(charge+1) (r^2 + 1)^(-1/2) (:452-454) with that Ip (:457);
    * the WFAT sampler reads IonPotential(mol, ridx) = -E(HOMO + ridx)
      (:444-449, :148-159) and the WFAT integrals wfat_intdata[ridx] and
      dipole wfat_mu[ridx] (:228-233). Closed shell: ridx is an integer;
      open shell (two rows of energies): ridx = (spin, ridx) with spin 1 or 2.
    """

    def __init__(self, data: dict, rot_α=0.0, rot_β=0.0, rot_γ=0.0, soft_core=1.0):
        self.data = data
        self.name = data["name"]
        self.charge = int(data["charge"])
        self.nucl_charge = self.charge + 1
        E = np.asarray(data["energy_levels"], dtype=np.float64)
        occ = np.asarray(data["orbit_occ"], dtype=np.float64)
        self.open_shell = E.ndim == 2
        self.energy_levels = E if self.open_shell else E[None, :]
        occ = occ if self.open_shell else occ[None, :]
        self.homo_idx = []
        for row in occ:                                  # _get_HOMO_idx (:192-195), 0-based
            nz = np.nonzero(row)[0]
            if nz.size == 0:
                raise ValueError("[GenericMolecule] no occupied orbital")
            self.homo_idx.append(int(nz[-1]))
        self.Ip = -max(float(self.energy_levels[k, h]) for k, h in enumerate(self.homo_idx))
        self.rot = (float(rot_α), float(rot_β), float(rot_γ))
        self.soft_core = float(soft_core)
        self.l = self.m = 0
        self.asymp_coeff = 1.0
        self.quan_ax_θ = self.quan_ax_ϕ = 0.0

    def _key(self, orbit_ridx):
        if self.open_shell:
            if not (isinstance(orbit_ridx, tuple) and len(orbit_ridx) == 2 and orbit_ridx[0] in (1, 2)):
                raise ValueError("[GenericMolecule] For an open-shell molecule, `orbit_ridx` should be a two-element "
                                 "tuple `(spin, ridx)`.")
            return (int(orbit_ridx[0]), int(orbit_ridx[1]))
        if isinstance(orbit_ridx, tuple) or int(orbit_ridx) != orbit_ridx:
            raise ValueError("[GenericMolecule] For a closed-shell molecule, `orbit_ridx` should be an integer.")
        return int(orbit_ridx)

    def MolEnergyLevel(self, orbit_ridx) -> float:
        """GenericMolecule.jl:148-159."""
        k = self._key(orbit_ridx)
        spin, r = (k[0] - 1, k[1]) if self.open_shell else (0, k)
        i = self.homo_idx[spin] + r
        if not 0 <= i < self.energy_levels.shape[1]:
            raise ValueError(f"[GenericMolecule] orbit_ridx {orbit_ridx} is out of range")
        return float(self.energy_levels[spin, i])

    def ion_potential(self, orbit_ridx=0) -> float:
        return -self.MolEnergyLevel(orbit_ridx)

    def wfat_data(self, orbit_ridx=0):
        """(mu, intdata) of MolWFATData (:228-233): the orbital's dipole and the
        complex array intdata[nxi, m + m_max, l, m' + l] (0-based)."""
        key = self._key(orbit_ridx)
        norm = (lambda k: tuple(int(v) for v in k)) if self.open_shell else int
        table = {norm(k): v for k, v in self.data.get("wfat_intdata", {}).items()}
        mus = {norm(k): v for k, v in self.data.get("wfat_mu", {}).items()}
        if key not in table or key not in mus:
            raise ValueError(f"[GenericMolecule] The WFAT data of orbit_ridx={orbit_ridx} is not available "
                             "(the PySCF calculation is not part of this engine).")
        return np.asarray(mus[key], dtype=np.float64), np.asarray(table[key], dtype=np.complex128)

    def MolAsympCoeff(self, orbit_ridx=0) -> np.ndarray:
        """GenericMolecule.jl:304-309: the orbital's asymptotic coefficients,
        complex C[l, m + l] (0-based), l = 0..l_max."""
        key = self._key(orbit_ridx)
        norm = (lambda k: tuple(int(v) for v in k)) if self.open_shell else int
        table = {norm(k): v for k, v in self.data.get("asymp_coeff", {}).items()}
        if key not in table:
            raise ValueError(f"[GenericMolecule] The asymptotic coefficients of orbit_ridx={orbit_ridx} are not "
                             "available (the PySCF calculation is not part of this engine).")
        return np.asarray(table[key], dtype=np.complex128)

    def _desc(self) -> _abi.TargetDesc:
        d = _abi.TargetDesc()
        d.kind = _abi.TARGET_HYDROGEN_LIKE
        d.nucl_charge = self.nucl_charge
        d.Ip = self.Ip
        d.asymp_coeff = 1.0
        d.soft_core = self.soft_core
        return d

    def with_rotation(self, rot_α, rot_β, rot_γ) -> "GenericMolecule":
        return GenericMolecule(self.data, rot_α, rot_β, rot_γ, self.soft_core)

    def __eq__(self, o):
        return (type(o) is type(self) and self.data is o.data and self.rot == o.rot
                and self.soft_core == o.soft_core)

    def __repr__(self):
        return f"[GenericMolecule] {self.name}, Ip={self.Ip:.4f}, Z={self.nucl_charge}, rot={self.rot}"


_MOLECULE_FILES = {   # MoleculeDatabase.jl:54-68
    "Hydrogen": "Molecule_Hydrogen.jld2", "Nitrogen": "Molecule_Nitrogen.jld2", "Oxygen": "Molecule_Oxygen.jld2",
    "Carbon Monoxide": "Molecule_CarbonMonoxide.jld2", "Nitric Oxide": "Molecule_NitricOxide.jld2",
    "Hydrochloric Acid": "Molecule_HydrochloricAcid.jld2", "Carbon Dioxide": "Molecule_CarbonDioxide.jld2",
    "Sulfur Dioxide": "Molecule_SulfurDioxide.jld2", "Water": "Molecule_Water.jld2",
    "Ammonia": "Molecule_Ammonia.jld2", "Acetylene": "Molecule_Acetylene.jld2", "Methane": "Molecule_Methane.jld2",
    "Benzene": "Molecule_Benzene.jld2",
}


def get_available_mols():
    """MoleculeDatabase.jl:50-52."""
    return list(_MOLECULE_FILES)


def get_mol(name: str, rot_α=0.0, rot_β=0.0, rot_γ=0.0, database_dir=None) -> GenericMolecule:
    """MoleculeDatabase.jl:42-48: a database molecule, read from eTraj.jl's
    `src/Targets/Database` directory (`database_dir`, or the environment
    variable ETRAJ_MOL_DATABASE)."""
    import os
    from .moldata import read_molecule_jld2
    if name not in _MOLECULE_FILES:
        raise ValueError(f"[get_mol] Molecule name `{name}` not found in database.")
    d = database_dir or os.environ.get("ETRAJ_MOL_DATABASE")
    if not d:
        raise FileNotFoundError("[get_mol] set database_dir or ETRAJ_MOL_DATABASE to eTraj.jl's src/Targets/Database")
    return GenericMolecule(read_molecule_jld2(os.path.join(d, _MOLECULE_FILES[name])), rot_α, rot_β, rot_γ)


def MolRotation(mol: MoleculeBase):
    """GenericMolecule.jl: (rot_α, rot_β, rot_γ) in radians."""
    return mol.rot


def synthetic_g00_table(seed: int, na: int = 8, nb: int = 16, mean: float = 1.0, sigma: float = 0.3) -> np.ndarray:
    """A seeded G00 coefficient table c_ab ~ mean + sigma N(0,1) (real and imaginary parts independent, the
    imaginary part centred on 0), for workloads that stand in for a quantum-chemistry structure factor."""
    rng = np.random.default_rng(seed)
    return (mean + sigma * rng.standard_normal((na, nb))) + 1j * sigma * rng.standard_normal((na, nb))


# ---- accessors (Targets/exports.jl) ------------------------------------------
def IonPotential(t): return t.Ip
def AsympNuclCharge(t): return t.nucl_charge
def AngularQuantumNumber(t): return t.l
def MagneticQuantumNumber(t): return t.m
def AsympCoeff(t): return t.asymp_coeff
def SoftCore(t): return t.soft_core
def TargetName(t): return t.name
def QuantizationAxisOrientaion(t): return (t.quan_ax_θ, t.quan_ax_ϕ)


def _eval(t, pts):
    ctx = _abi.default_context()
    xyz = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    n = xyz.shape[0]
    V = np.empty(n)
    F = np.empty((n, 3))
    ctx.check(ctx.lib.etraj_target_eval(ctx.handle, t._desc(), _abi.dptr(xyz), n, _abi.dptr(V), _abi.dptr(F)))
    return V, F


def TargetPotential(t):
    """(x,y,z) -> V, device evaluation (HydrogenLikeAtom.jl:75, SAEAtom.jl:84)."""
    def V(x, y, z):
        v, _ = _eval(t, np.stack(np.broadcast_arrays(x, y, z), axis=-1))
        return v if np.ndim(x) or np.ndim(y) or np.ndim(z) else float(v[0])
    return V


def TargetForce(t):
    """(x,y,z) -> (Fx,Fy,Fz), device evaluation (HydrogenLikeAtom.jl:76, SAEAtom.jl:91)."""
    def F(x, y, z):
        _, f = _eval(t, np.stack(np.broadcast_arrays(x, y, z), axis=-1))
        return tuple(f[:, k] for k in range(3)) if np.ndim(x) or np.ndim(y) or np.ndim(z) else tuple(f[0])
    return F


def Serialize(t) -> dict:
    """Targets' Serialize (HydrogenLikeAtom.jl, SAEAtom.jl, GenericMolecule.jl:620-633)."""
    d = {"type": type(t).__name__}
    if isinstance(t, GenericMolecule):
        d.update(atoms=list(t.data.get("atoms", [])), atom_coords=np.asarray(t.data.get("atom_coords", [])).tolist(),
                 charge=t.charge, spin=int(t.data.get("spin", 0)), name=t.name, rot_α=t.rot[0], rot_β=t.rot[1],
                 rot_γ=t.rot[2])
        return d
    d.update({k: v for k, v in t.__dict__.items() if not isinstance(v, np.ndarray)})
    return d


# ---- AtomDatabase.jl:54-275 -------------------------------------------------
_HYD = {"H": (0.5, 1, "H"), "He1p": (1.0, 2, "He⁺"), "Li2p": (1.5, 3, "Li²⁺")}
# key: (Ip, Z, a1, b1, a2, b2, a3, b3, l, name)
_SAE = {
    "He": (0.9035698802, 1, 1.230723, 0.6620055, -1.325040, 1.236224, -0.2307230, 0.4804286, 0, "He"),
    "Ne": (0.79248225, 1, 8.069321, 2.148149, -3.569610, 1.985509, 0.9306791, 0.6018819, 1, "Ne"),
    "Ne1p": (1.505361, 2, 8.042896, 2.714635, 0.5064300, 0.9818268, -0.04289633, 0.4012269, 1, "Ne⁺"),
    "Ne2p": (2.3307635, 3, 7.000000, 3.007574, 0.6521344, 1.048617, 0.0, 0.0, 1, "Ne²⁺"),
    "Ar": (0.579155055, 1, 16.03862, 2.007423, -25.54278, 4.524663, 0.9613848, 0.4426447, 1, "Ar"),
    "Ar1p": (1.0153715, 2, 14.98869, 2.217293, -23.60558, 4.584761, 1.011314, 0.5510695, 1, "Ar⁺"),
    "Ar2p": (1.497, 3, 15.00000, 3.588835, 5.951071, 1.692914, 0.0, 0.0, 1, "Ar²⁺"),
    "V": (0.2479178, 1, 21.27861, 1.810071, -30.02658, 3.201500, 0.7213853, 0.4534235, 2, "V"),
    "Ni": (0.28076035, 1, 26.81231, 1.937326, -39.55201, 3.161884, 0.1876890, 0.2626299, 2, "Ni"),
    "Kr": (1.02895202, 1, 23.6188, 10.8391, 109.956, 6.55432, 11.8312, 1.27723, 1, "Kr"),
    "Kr1p": (1.790416, 2, 23.1460, 11.0416, 110.447, 6.67063, 10.8540, 1.41524, 1, "Kr⁺"),
    "Rb": (0.153506625, 1, 24.02321, 11.10663, 115.1998, 6.629123, 11.97679, 1.244882, 0, "Rb"),
    "Nb": (0.248383, 1, 4.000000, 1.379674, -6.356689, 2.245839, 0.0, 0.0, 2, "Nb"),
    "Pd": (0.3063732, 1, 43.83606, 2.395037, -93.35653, 4.907321, 1.163943, 0.3767380, 2, "Pd"),
    "Xe": (0.445763535, 1, 51.35554, 2.111554, -99.92747, 3.737221, 1.644457, 0.4306465, 1, "Xe"),
    "Xe1p": (0.7708, 2, 37.49194, 9.452474, 103.1967, 5.305537, 14.50806, 1.315683, 1, "Xe⁺"),
    "Ta": (0.27744165, 1, 67.70192, 2.565180, -135.5191, 4.088800, 4.298078, 0.7303277, 2, "Ta"),
}


def get_available_atoms():
    return list(_HYD) + list(_SAE)


def get_atom(name: str, charge: int = 0, **kwargs):
    """AtomDatabase.jl:33-47."""
    key = name + (f"{charge}p" if charge > 0 else f"{-charge}n" if charge < 0 else "")
    if key in _HYD:
        Ip, Z, nm = _HYD[key]
        args = dict(Ip=Ip, Z=Z, name=nm)
        args.update(kwargs)
        return HydrogenLikeAtom(**args)
    if key in _SAE:
        Ip, Z, a1, b1, a2, b2, a3, b3, l, nm = _SAE[key]
        args = dict(Ip=Ip, Z=Z, a1=a1, b1=b1, a2=a2, b2=b2, a3=a3, b3=b3, l=l, name=nm)
        args.update(kwargs)
        return SAEAtom(**args)
    raise KeyError(f"[get_atom] Atom name key `{key}` not found in database.")
