"""The experiment harness on the GPU: a small coupled double-gyre experiment staged from
CDL files and run twice in a chain by openqg_amd.run_model (the flow of the reference's
run_model.py), the second run starting from the first one's lastday files."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from experiment_files import write_experiment
from openqg_amd import run_model as RM
from openqg_amd.cdl import read_cdl

pytestmark = pytest.mark.gpu


def _read(path, names):
    with netcdf_file(path, "r", mmap=False) as f:
        return [np.array(f.variables[n][:]) for n in names]


def test_chained_coupled_experiment(gpu, tmp_path):
    src = write_experiment(str(tmp_path / "input"), days=0.5, resday=0.25)
    out = str(tmp_path / "run")
    print("[experiment] chained coupled runs", flush=True)
    RM.main(out, repeats=2, input_dir=src)
    for r in (0, 1):
        o = os.path.join(out, f"out-{r}")
        files = sorted(os.listdir(o))
        for n in ("small_ocean_qg_lastday.nc", "small_atmos_qg_lastday.nc", "small_ocean_ml_lastday.nc",
                  "small_atmos_ml_lastday.nc", "small_ocean_qg_restart.nc", "stdout.txt"):
            assert n in files, (r, n, files)
        p, pm = _read(os.path.join(o, "small_ocean_qg_lastday.nc"), ("p", "pm"))
        assert p.shape == (3, 33, 33) and np.isfinite(p).all() and np.abs(p).max() > 0
        pa, = _read(os.path.join(o, "small_atmos_qg_lastday.nc"), ("p",))
        assert pa.shape == (3, 9, 17) and np.isfinite(pa).all()
        sst, sstm = _read(os.path.join(o, "small_ocean_ml_lastday.nc"), ("sst", "sstm"))
        ast, hmixa = _read(os.path.join(o, "small_atmos_ml_lastday.nc"), ("ast", "hmixa"))
        assert sst.shape == (32, 32) and ast.shape == (8, 16) and np.isfinite(sst).all() and np.isfinite(hmixa).all()
        assert 50.0 <= hmixa.min() and hmixa.max() <= 5000.0
    # the second run started where the first ended
    c1 = read_cdl(os.path.join(out, "in-1", "clock.cdl"))
    assert abs(c1["tini"] - 0.5 / 365.0) <= 1e-5
    o1 = read_cdl(os.path.join(out, "in-1", "ocn_basin.cdl"))
    assert o1["qg_state"] == "small_ocean_qg_lastday.nc" and o1["oml_state"] == "small_ocean_ml_lastday.nc"
    a = _read(os.path.join(out, "out-0", "small_ocean_qg_lastday.nc"), ("p",))[0]
    b = _read(os.path.join(out, "in-1", "small_ocean_qg_lastday.nc"), ("p",))[0]
    np.testing.assert_array_equal(a, b)
    log = open(os.path.join(out, "out-1", "stdout.txt")).read()
    assert "steps 145..288" in log and "done at step 288" in log
