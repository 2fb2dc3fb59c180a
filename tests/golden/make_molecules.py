"""Writes tests/golden/mol_*.npz: the fields of four molecule files of
eTraj.jl's database (src/Targets/Database/Molecule_*.jld2, read with
etraj_hip.moldata) that the WFAT path uses. Run where the reference is
present:  python tests/golden/make_molecules.py /root/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "etraj.jl_amd"))
from etraj_hip.moldata import read_molecule_jld2  # noqa: E402

MOLS = {"Hydrogen": "Molecule_Hydrogen.jld2", "CarbonMonoxide": "Molecule_CarbonMonoxide.jld2",
        "Water": "Molecule_Water.jld2", "Oxygen": "Molecule_Oxygen.jld2"}


def main(ref):
    db = os.path.join(ref, "src", "Targets", "Database")
    for key, fn in MOLS.items():
        m = read_molecule_jld2(os.path.join(db, fn))
        out = {k: np.asarray(m[k]) for k in ("name", "charge", "spin", "energy_levels", "orbit_occ")}
        keys = sorted(m["wfat_intdata"], key=str)
        out["wfat_keys"] = np.array([str(k) for k in keys])
        for i, k in enumerate(keys):
            out[f"wfat_intdata_{i}"] = m["wfat_intdata"][k]
            out[f"wfat_mu_{i}"] = np.asarray(m["wfat_mu"][k])
        akeys = sorted(m["asymp_coeff"], key=str)
        out["asymp_keys"] = np.array([str(k) for k in akeys])
        for i, k in enumerate(akeys):
            out[f"asymp_coeff_{i}"] = m["asymp_coeff"][k]
        np.savez_compressed(os.path.join(HERE, f"mol_{key}.npz"), **out)
        print(key, keys)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
