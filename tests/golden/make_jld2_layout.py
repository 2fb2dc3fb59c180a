"""Writes tests/golden/jld2_layout.json: the container layout (user block and its
signature, superblock version, root-group link storage, String encoding; read
with etraj_hip.h5out.file_layout) of JLD2 files that ship with eTraj.jl
(src/Targets/Database/Molecule_*.jld2, written by JLD2.jl). The JLD2 output
test holds write_jld2's files to it. Run where the reference is present:
    python tests/golden/make_jld2_layout.py /root/reference
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "etraj.jl_amd"))
from etraj_hip.h5out import file_layout  # noqa: E402

FILES = ["Molecule_Hydrogen.jld2", "Molecule_Water.jld2", "Molecule_CarbonMonoxide.jld2"]


def main(ref):
    db = os.path.join(ref, "src", "Targets", "Database")
    out = {fn: file_layout(os.path.join(db, fn), "name") for fn in FILES}
    with open(os.path.join(HERE, "jld2_layout.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
