"""Output files of perform_traj_simulation (write_output!, TrajectorySimulation.jl:
502-546, format :h5): dataset names, types, Julia's reversed dimension order, the
shuffle + deflate filters of output_compress, the temp-file rename and the file
naming rules (:195-215). Read back with the HDF5 C library (etraj_hip.moldata)
and, where the HDF5 tools are installed, with h5dump."""
import os
import shutil
import subprocess
import warnings

import numpy as np
import pytest

from etraj_hip import Cos2Laser, get_atom
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


@pytest.mark.parametrize("dim,compress", [(2, True), (2, False), (3, True)])
def test_h5_file_has_the_reference_datasets(tmp_path, dim, compress):
    r, spec, num, pmax = _fake_result(dim)
    path = str(tmp_path / "run")
    _save(r, path, "h5", compress)
    assert r["output_path"] == path + ".h5" and os.path.exists(path + ".h5") and not os.path.exists(path + ".h5!")
    f = _File(path + ".h5")
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
        out = subprocess.run([H5DUMP, "-H", "-p", "-d", "/momentum_spec", path + ".h5"], capture_output=True,
                             text=True, cwd=str(tmp_path)).stdout
        assert f"( {', '.join(str(n) for n in num[::-1])} )" in out    # HDF5 dims = Julia's reversed
        assert ("SHUFFLE" in out and "DEFLATE" in out and "LEVEL 6" in out) == compress
        out = subprocess.run([H5DUMP, "-H", "-d", "/num_effective_traj", path + ".h5"], capture_output=True,
                             text=True, cwd=str(tmp_path)).stdout
        assert "H5T_STD_I64LE" in out and "SCALAR" in out


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


def test_jld2_and_unknown_formats_are_refused(tmp_path):
    r, *_ = _fake_result()
    with pytest.raises(NotImplementedError):
        _save(r, str(tmp_path / "x"), "jld2")
    with pytest.raises(ValueError):
        _save(r, str(tmp_path / "x"), "csv")
