"""etraj_hip -- the HIP/gfx950 trajectory engine behind eTraj.jl's
perform_traj_simulation, with the reference's Lasers / Targets /
ElectronSamplers interface.

Every compute entry point calls libetraj_hip.so (include/etraj.h); there is
no CPU fallback.
"""
__version__ = "0.1.0"

from . import _abi
from ._abi import Context, EngineError, EngineUnavailable, load_library
from .lasers import (Laser, MonochromaticLaser, Cos2Laser, Cos4Laser, GaussianLaser, BichromaticLaser, LaserFx,
                     LaserFy, LaserAx, LaserAy, LaserF0, LaserA0, PeakInt, WaveLen, CycNum, SpreadCycNum, Ellipticity,
                     Azimuth, AngFreq, Period, CEP, TimeShift, KeldyshParameter, SpreadDuration, FWHM_Duration, Laser1,
                     Laser2, Delay21)
from .targets import (Target, SAEAtomBase, HydrogenLikeAtom, SAEAtom, MoleculeBase, TabulatedMolecule, MolRotation,
                      GenericMolecule, get_mol, get_available_mols,
                      synthetic_g00_table, get_atom, get_available_atoms,
                      hartree_asymp_coeff, IonPotential, AsympNuclCharge, AngularQuantumNumber, MagneticQuantumNumber,
                      AsympCoeff, SoftCore, TargetName, QuantizationAxisOrientaion, TargetPotential, TargetForce)
from .samplers import ElectronSampler, ADKSampler, SPASampler, SPANESampler, WFATSampler, init_sampler, batch_num, batch_max_size, gen_electron_batch
from .simulation import perform_traj_simulation, perform_traj_scan, run_desc, autotune

__all__ = [n for n in dir() if not n.startswith("_")]
