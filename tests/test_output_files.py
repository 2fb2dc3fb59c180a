"""Output files of perform_traj_simulation (write_output!, TrajectorySimulation.jl:
502-546, formats :h5 and :jld2): dataset names, types, Julia's reversed dimension
order, the filters of output_compress (shuffle + deflate for :h5, deflate for
:jld2), the temp-file rename and the file naming rules (:195-215). Read back with
the HDF5 C library through etraj_hip.moldata, the reader of the reference's own
.jld2 database files, and, where the HDF5 tools are installed, with h5dump. A JLD2
file's container layout (user block and signature, superblock version, root-group
link storage, String encoding) must equal that of the JLD2 files that ship with
eTraj.jl (tests/golden/jld2_layout.json, written by make_jld2_layout.py)."""
import json
import os
import shutil
import subprocess
import warnings

import numpy as np
import pytest

from etraj_hip import Cos2Laser, get_atom
from etraj_hip.h5out import file_layout
from etraj_hip.moldata import _File, HDF5Unavailable, _hdf5
from etraj_hip.simulation import _result, _save

try:
    _hdf5()
    HAVE_HDF5 = True
except HDF5Unavailable:
    HAVE_HDF5 = False
pytestmark = pytest.mark.skipif(not HAVE_HDF5, reason="libhdf5 not found")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _fake_result(dim=2):
    num = (7, 5, 3)[:dim]
    pmax = (1.5, 1.0, 0.5)[:dim]
    spec = np.arange(int(np.prod(num)), dtype=float) * 0.25       # px fastest, as the engine returns
    kw = dict(init_cond_method="ADK", laser=Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0),
              target=get_atom("H"), dimension=dim, sample_t_intv=(-100.0, 100.0), sample_t_num=10,
              traj_phase_method="CTMC", traj_t_final=120.0, final_p_max=pmax, final_p_num=num,
              traj_integrator="rk4", traj_dt=0.1, precision="fp64")
    return _result(spec, 0.0125, 0.0025, 12345, {}, **kw), spec, num, pmax


@pytest.mark.parametrize("fmt", ["h5", "jld2"])
@pytest.mark.parametrize("dim,compress", [(2, True), (2, False), (3, True)])
def test_output_file_has_the_reference_datasets(tmp_path, fmt, dim, compress):
    r, spec, num, pmax = _fake_result(dim)
    path = str(tmp_path / "run")
    ext = "." + fmt
    _save(r, path, fmt, compress)
    assert r["output_path"] == path + ext and os.path.exists(path + ext) and not os.path.exists(path + ext + "!")
    f = _File(path + ext)
    try:
        ms = f.get("momentum_spec")
        assert ms.shape == num                                       # Julia's size, read back in Julia order
        np.testing.assert_array_equal(ms.ravel(order="F"), spec)     # px fastest
        for ax, name in enumerate(("px", "py", "pz")[:dim]):
            np.testing.assert_array_equal(f.get(name), np.linspace(-pmax[ax], pmax[ax], num[ax]))
        assert f.get("ion_prob") == 0.0125 and f.get("ion_prob_uncollected") == 0.0025
        assert f.get("num_effective_traj") == 12345
        text = f.get("params_text")
        assert "init_cond_method: ADK" in text and "Cos2Laser" in text
        assert f.get("info").startswith("Trajectory simulation output generated by")
    finally:
        f.close()
    if H5DUMP:
        out = subprocess.run([H5DUMP, "-H", "-p", "-d", "/momentum_spec", path + ext], capture_output=True,
                             text=True, cwd=str(tmp_path)).stdout
        assert f"( {', '.join(str(n) for n in num[::-1])} )" in out    # HDF5 dims = Julia's reversed
        assert ("DEFLATE" in out and "LEVEL 6" in out) == compress
        assert ("SHUFFLE" in out) == (compress and fmt == "h5")
        out = subprocess.run([H5DUMP, "-H", "-d", "/num_effective_traj", path + ext], capture_output=True,
                             text=True, cwd=str(tmp_path)).stdout
        assert "H5T_STD_I64LE" in out and "SCALAR" in out


def test_jld2_container_matches_the_reference_files(tmp_path):
    r, *_ = _fake_result(3)
    _save(r, str(tmp_path / "run"), "jld2")
    got = file_layout(r["output_path"], "info")
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "jld2_layout.json")))
    assert len(ref) >= 3
    for fn, want in ref.items():
        assert got == want, fn
    head = open(r["output_path"], "rb").read(512)
    assert head.startswith(b"HDF5-based Julia Data Format, version 0.1.1\x00")
    assert head.rstrip(b"\x00").count(b"\x00") == 1          # signature, then the writer's note
    # the default format is the reference's (output_fmt = :jld2, TrajectorySimulation.jl:127)
    r2, *_ = _fake_result()
    _save(r2, str(tmp_path / "default"))
    assert r2["output_path"].endswith(".jld2")


def test_existing_files_are_not_overwritten(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    r, *_ = _fake_result()
    _save(r, "a", "h5")
    first = r["output_path"]
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        _save(r, "a", "h5")
    assert r["output_path"] != first and r["output_path"].endswith(".h5")
    assert any("already exists" in str(x.message) for x in w)
    open("b.h5!", "w").close()                                       # a temp file of an earlier session
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        _save(r, "b.h5", "hdf5")
    assert r["output_path"] != "b.h5" and any("temp file" in str(x.message) for x in w)


def test_unknown_formats_are_refused(tmp_path):
    r, *_ = _fake_result()
    with pytest.raises(ValueError):
        _save(r, str(tmp_path / "x"), "csv")
