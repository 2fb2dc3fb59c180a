"""Descriptor builders shared by the tests: one keyword set -> the C-ABI
descriptors that drive both the engine and the oracle."""
import warnings

import numpy as np

from etraj_hip import _abi, workloads
from etraj_hip.samplers import init_sampler
from etraj_hip.simulation import run_desc


def descriptors(kw):
    kw = dict(kw)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(**kw)
    rd = run_desc(**kw)
    return kw["laser"]._desc(), kw["target"]._desc(), sp.desc, rd


def small_ctmc(n_t=40, n_kd=24, precision="fp64", **over):
    kw = workloads.h_lp_ctmc_grid(n_t=n_t, n_kd=n_kd, precision=precision)
    kw.update(over)
    return kw


def small_scts(n_t=24, n_k=96, precision="fp64", p_num=200, **over):
    kw = workloads.h_lp_scts_stratified(n_t=n_t, n_k=n_k, precision=precision, p_num=p_num)
    kw.update(over)
    return kw


def l1(p, q):
    p = np.asarray(p, float).ravel()
    q = np.asarray(q, float).ravel()
    return float(np.abs(p - q).sum() / np.abs(q).sum())


def small_wfat(n_t=24, n_k=96, precision="fp64", p_num=160, **over):
    kw = workloads.mol_lp_scts_wfat(n_t=n_t, n_k=n_k, precision=precision, p_num=p_num, rot=(0.3, 1.1, -0.7))
    kw.update(over)
    return kw


def fixture_molecule(key, rot=(0.0, 0.0, 0.0)):
    """GenericMolecule from tests/golden/mol_<key>.npz (fields of eTraj.jl's
    database file, written by tests/golden/make_molecules.py)."""
    import ast
    import os
    from etraj_hip.targets import GenericMolecule
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", f"mol_{key}.npz"))
    data = {k: f[k] for k in ("energy_levels", "orbit_occ")}
    data.update(name=str(f["name"]), charge=int(f["charge"]), spin=int(f["spin"]), wfat_intdata={}, wfat_mu={})
    for i, k in enumerate(f["wfat_keys"]):
        k = ast.literal_eval(str(k))
        data["wfat_intdata"][k] = f[f"wfat_intdata_{i}"]
        data["wfat_mu"][k] = f[f"wfat_mu_{i}"]
    data["asymp_coeff"] = {ast.literal_eval(str(k)): f[f"asymp_coeff_{i}"] for i, k in enumerate(f["asymp_keys"])}
    return GenericMolecule(data, *rot)
