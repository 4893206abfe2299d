"""The QG-only run loop (openqg_amd/driver.py) on the GPU: an ocean box and an
atmosphere channel advanced together for two days at the coupled step ratio
(dta = 200 s, nstr = 18), against the same two basins stepped one step at a
time with the reference's per-step cadence (src/openqg.F90:483-509).  The
states must agree bit for bit (graph replay and single steps run the same
kernels in the same order), and the restart files must hold them.
"""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from openqg_amd import workloads as W
from openqg_amd.driver import Basin, Clock, Valids, run

pytestmark = pytest.mark.gpu


def _make(w):
    qg = w.make_qg()
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    qg.set_forcing(*w.forcing())
    return qg


def test_run_loop_matches_single_steps(gpu, tmp_path):
    wo, wa = W.small_box(), W.small_channel()
    clk = Clock(dta=200.0, nstr=18, tini=0.0, trun=2.0 / 365.0, resday=0.5)
    assert clk.dto == wo.dt and clk.dta == wa.dt
    print(f"[driver] {wo.name} + {wa.name}, {clk.ntsrun} steps", flush=True)
    o1, a1 = _make(wo), _make(wa)
    seen, samples = [], []
    nt = run(clk, str(tmp_path), ocn=Basin("ocean", o1, Valids(100, 1e12, 1e3), nocmon=144),
             atm=Basin("atmos", a1, Valids(100, 1e12, 1e3)),
             monitor=lambda n, t, d: (seen.append((n, t)), samples.append(d)))
    assert nt == clk.ntsrun == 864
    assert seen == [("ocean", t) for t in range(0, 865, 144)]
    # each sample carries the energetics and pket from the previous sample's et2m
    for k, d in enumerate(samples):
        assert {"ekei", "ah4d", "btdg", "utau", "pket"} <= set(d)
        assert np.all(np.isfinite(d["ah4d"])) and d["btdg"] > 0.0
        if k == 0:
            assert np.all(d["pket"] == 0.0)
        else:
            dt = 144 * clk.dta
            want = 0.5 * o1.rho * np.asarray(o1.gp) * (d["et2m"] - samples[k - 1]["et2m"]) / dt
            np.testing.assert_allclose(d["pket"], want, rtol=1e-14, atol=0)

    o2, a2 = _make(wo), _make(wa)
    for t in range(1, clk.nsteps + 1):
        a2.step(clk.tdta)
        if (t - 1) % 100 == 0:
            a2.suppress()
        if (t - 1) % clk.nstr == 0:
            o2.step(clk.tdto)
            if (t - 1) % (25 * clk.nstr) == 0:
                o2.suppress()
    for name, x, y in (("ocean", o1, o2), ("atmos", a1, a2)):
        sx, sy = x.get_state(), y.get_state()
        for f, u, v in zip(("q", "qm", "p", "pm"), sx, sy):
            assert u.tobytes() == v.tobytes(), f"{name} {f}"
        np.testing.assert_array_equal(x.get_constr(), y.get_constr())
        assert np.isfinite(sx[2]).all()
        for suffix in ("lastday", "restart"):
            with netcdf_file(os.path.join(tmp_path, f"{name}_qg_{suffix}.nc"), "r", mmap=False) as f:
                assert np.array(f.variables["p"][:], dtype=np.float64).tobytes() == sx[2].tobytes()
                assert np.array(f.variables["pm"][:], dtype=np.float64).tobytes() == sx[3].tobytes()
