"""The experiment harness on the CPU: CDL configuration files read as text, the mesh
and boxes built from them (mesh.F90, box.F90:50-90), and the chaining edits of
run_model.py (update_clock, update_basin)."""
import math
import os

import pytest

from experiment_files import write_experiment
from openqg_amd import run_model as RM
from openqg_amd.cdl import parse_cdl, read_cdl
from openqg_amd.coupled import load_grid
from openqg_amd.driver import Clock
from openqg_amd.experiment import OMEGA, Mesh, box_from_mesh


def test_parse_cdl_types_arrays_strings_comments():
    d = parse_cdl('''netcdf x {
dimensions:
    strlen = 64 ; nl = 3 ;
variables:
    double a; integer n; double h(nl); char s(strlen); double e, f;
data:
    a = 1.5e3; // a comment; with a semicolon
    n = 7;
    h = 300.0,  1100.0, 2600.0;
    s = "with ; semicolon";
    e = 2.0d-3; f = -4;
}''')
    assert d["a"] == 1500.0 and d["n"] == 7 and isinstance(d["n"], int)
    assert d["h"] == [300.0, 1100.0, 2600.0]
    assert d["s"] == "with ; semicolon"
    assert d["e"] == 2.0e-3 and d["f"] == -4.0
    assert d["__dims__"] == {"strlen": 64, "nl": 3}


def test_mesh_and_boxes(tmp_path):
    write_experiment(str(tmp_path))
    g = read_cdl(str(tmp_path / "glam.cdl"))
    m = Mesh(g["nxt"], g["nyt"], g["dx"], g["dy"], g["x0"], g["y0"], bool(g["cyclic"]), g["x1"], g["lat"])
    phi = math.radians(43.29)
    assert math.isclose(m.fnot, 2 * OMEGA * math.sin(phi), rel_tol=1e-15)
    assert math.isclose(m.beta, 2 * OMEGA * math.cos(phi) / (6378.1e3 * math.cos(phi)), rel_tol=1e-15)
    go = box_from_mesh(m, read_cdl(str(tmp_path / "ocn_basin.cdl")), -1)
    ga = box_from_mesh(m, read_cdl(str(tmp_path / "atm_basin.cdl")), 1)
    assert (go.nxp, go.nyp, go.dx, go.cyclic, go.y0) == (33, 33, 15e3, False, 240e3)
    assert (ga.nxp, ga.nyp, ga.dx, ga.cyclic) == (17, 9, 120e3, True)
    g = load_grid(go, ga)
    # the ocean's own placement in the mesh and the coupling grid's centring agree
    assert (g.ndxr, g.ny1) == (8, 3) and go.y0 == (g.ny1 - 1) * ga.dy


def test_chaining_edits(tmp_path):
    write_experiment(str(tmp_path), days=10.0, tini=0.0, qg_o="zero", oml="rbal")
    c0 = read_cdl(str(tmp_path / "clock.cdl"))
    RM.update_clock(str(tmp_path / "clock.cdl"))
    c = read_cdl(str(tmp_path / "clock.cdl"))
    # written with six decimals as the reference's update_clock does; the next run starts where this one ended
    assert math.isclose(c["tini"], 10.0 / 365.0, rel_tol=1e-4) and math.isclose(c["trun"], 10.0 / 365.0, rel_tol=1e-4)
    k0 = Clock(c0["dta"], c0["nstr"], c0["tini"], c0["trun"], c0["resday"])
    k1 = Clock(c["dta"], c["nstr"], c["tini"], c["trun"], c["resday"])
    assert k1.nsteps0 == k0.nsteps and k1.ntsrun == k0.ntsrun
    RM.update_basin(str(tmp_path / "ocn_basin.cdl"))
    RM.update_basin(str(tmp_path / "atm_basin.cdl"))
    o, a = read_cdl(str(tmp_path / "ocn_basin.cdl")), read_cdl(str(tmp_path / "atm_basin.cdl"))
    assert o["qg_state"] == "small_ocean_qg_lastday.nc" and o["oml_state"] == "small_ocean_ml_lastday.nc"
    assert a["qg_state"] == "small_atmos_qg_lastday.nc" and a["aml_state"] == "small_atmos_ml_lastday.nc"


@pytest.mark.skipif(not os.path.isdir("/root/reference/data"), reason="the reference's data directory is absent")
def test_reference_configs_parse_and_place_the_ocean():
    base = "/root/reference/data"
    for sub in ("glam", "basin", "clocks", "windstress"):
        for fn in os.listdir(os.path.join(base, sub)):
            read_cdl(os.path.join(base, sub, fn))
    g = read_cdl(os.path.join(base, "glam", "glam-dg.cdl"))
    m = Mesh(g["nxt"], g["nyt"], g["dx"], g["dy"], g["x0"], g["y0"], bool(g["cyclic"]), g.get("x1", 0.0), g["lat"])
    go = box_from_mesh(m, read_cdl(os.path.join(base, "basin", "ocn_dg.cdl")), -1)
    ga = box_from_mesh(m, read_cdl(os.path.join(base, "basin", "atm_dg.cdl")), 1)
    gr = load_grid(go, ga)
    assert (go.nxp, go.nyp, ga.nxp, ga.nyp) == (385, 481, 129, 65)
    assert gr.ndxr == 12 and go.y0 == (gr.ny1 - 1) * ga.dy
