"""The run loop's schedule (openqg_amd/driver.py) against the reference's
per-step cadence, with stand-in basins that record what they are asked to do.

The reference decides at every atmosphere step nt (src/openqg.F90:483-509):
ocean_step = mod(nt-1, nstr) == 0, ocn_supp_step = mod(nt-1, 25 nstr) == 0,
atm_supp_step = mod(nt-1, 100) == 0, a restart dump when
mod(nt - nsteps0, noutre) == 0 and a validity check when
mod(nt - nsteps0, nvalid) == 0.  The driver batches steps between host events
into `QG.run` calls; expanding those calls must give back exactly the
per-step sequence.
"""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from openqg_amd.driver import Basin, Clock, InvalidSolution, Valids, nint, ocean_steps, run
from openqg_amd.qg import Box


class FakeQG:
    """Records steps; p holds the number of steps taken so far."""

    def __init__(self, name, bad_after=None):
        self.b = Box(5, 4, 2, 1.0e4, 1.0e4, 1.0e-4, 2.0e-11, [1000.0, 3000.0])
        self.name = name
        self.steps = []          # (tdt, suppressed) per step, in order
        self.diag_at = []        # number of steps taken at each diagnose call
        self.bad_after = bad_after

    def run(self, tdt, nsteps, step0=0, supp_every=25):
        assert nsteps >= 1
        for i in range(nsteps):
            self.steps.append((tdt, supp_every > 0 and (step0 + i) % supp_every == 0))

    def get_state(self):
        f = np.full((self.b.nl, self.b.nyp, self.b.nxp), float(len(self.steps)))
        return f, f, f, -f

    def diagnose(self, dt):
        n = len(self.steps)
        self.diag_at.append(n)
        big = 1e30 if self.bad_after is not None and n >= self.bad_after else 1.0
        z = np.zeros(self.b.nl)
        return dict(nonfinite=0.0, pmin=z - 1, pmax=z + big, qmin=z - 1, qmax=z + 1)


def _reference_schedule(clk):
    atm, ocn = [], []
    for nt in range(clk.nsteps0 + 1, clk.nsteps + 1):
        atm.append((clk.tdta, (nt - 1) % 100 == 0))
        if (nt - 1) % clk.nstr == 0:
            ocn.append((clk.tdto, (nt - 1) % (25 * clk.nstr) == 0))
    return atm, ocn


def _read_p(path):
    with netcdf_file(path, "r", mmap=False) as f:
        return float(f.variables["p"][0, 0, 0])


@pytest.mark.parametrize("dta,nstr,tini_days,days,resday", [
    (1200.0, 1, 0.0, 10.0, 0.0),      # ocean-only cadence, no intermediate dumps
    (200.0, 6, 0.0, 3.0, 1.0),        # coupled ratio, daily dumps
    (200.0, 6, 7.25, 2.5, 0.5),       # restart run starting mid-cycle
    (600.0, 7, 1.0 / 3.0, 4.0, 0.0),  # nstr not dividing the start step
])
def test_schedule_matches_per_step_cadence(tmp_path, dta, nstr, tini_days, days, resday):
    clk = Clock(dta=dta, nstr=nstr, tini=tini_days / 365.0, trun=days / 365.0, resday=resday)
    o, a = FakeQG("ocn"), FakeQG("atm")
    nt = run(clk, str(tmp_path), ocn=Basin("ocean", o, Valids(37, 1e9, 1e9)),
             atm=Basin("atmos", a, Valids(50, 1e9, 1e9)))
    assert nt == clk.nsteps
    ref_atm, ref_ocn = _reference_schedule(clk)
    assert a.steps == ref_atm
    assert o.steps == ref_ocn
    # validity checks at every nvalid-th step after the start, and at the start
    assert a.diag_at == [n for n in range(0, clk.ntsrun + 1) if n % 50 == 0]
    ocean_done = lambda nd: sum(1 for t in range(clk.nsteps0 + 1, clk.nsteps0 + nd + 1) if (t - 1) % nstr == 0)
    assert o.diag_at == [ocean_done(n) for n in range(0, clk.ntsrun + 1) if n % 37 == 0]
    # the last dump and the end of run hold the state of those steps
    assert _read_p(os.path.join(tmp_path, "atmos_qg_lastday.nc")) == len(ref_atm)
    assert _read_p(os.path.join(tmp_path, "ocean_qg_lastday.nc")) == len(ref_ocn)
    dumps = [n for n in range(1, clk.ntsrun + 1) if n % clk.noutre == 0]
    restart = os.path.join(tmp_path, "atmos_qg_restart.nc")
    if dumps:
        assert _read_p(restart) == dumps[-1]
        assert _read_p(os.path.join(tmp_path, "ocean_qg_restart.nc")) == ocean_done(dumps[-1])
    else:
        assert not os.path.exists(restart)


def test_clock_derived_values():
    clk = Clock(dta=200.0, nstr=6, tini=0.5, trun=30.0 / 365.0, resday=5.0)
    assert clk.dto == 1200.0 and clk.tdto == 2400.0 and clk.tdta == 400.0
    assert clk.nsteps0 == nint(0.5 * 365 * 86400 / 200.0) == 78840
    assert clk.ntsrun == 30 * 432 and clk.noutre == 5 * 432
    assert Clock(dta=1200.0, nstr=1, tini=0.0, trun=1.0 / 365.0).noutre == 72 + 1
    assert nint(2.5) == 3 and nint(-2.5) == -3 and nint(2.4999) == 2


def test_ocean_steps_counts():
    for nstr in (1, 3, 6):
        for nt0 in range(0, 20):
            for nt1 in range(nt0 + 1, 40):
                idx = [(nt - 1) // nstr for nt in range(nt0 + 1, nt1 + 1) if (nt - 1) % nstr == 0]
                first, count = ocean_steps(nt0, nt1, nstr)
                assert count == len(idx)
                if idx:
                    assert idx == list(range(first, first + count))


def test_invalid_solution_stops_and_saves(tmp_path):
    clk = Clock(dta=1200.0, nstr=1, tini=0.0, trun=5.0 / 365.0, resday=1.0)
    o = FakeQG("ocn", bad_after=100)
    seen = []
    with pytest.raises(InvalidSolution, match="nt = 108"):
        run(clk, str(tmp_path), ocn=Basin("ocean", o, Valids(12, 1e9, 1e9), nocmon=50),
            monitor=lambda name, nt, d: seen.append((name, nt, d["pmax"][0])))
    # the monitor's own cadence (0, 50, 100), then the sample the failed check
    # forces at step 108, of the state that failed
    assert [(n, t) for n, t, _ in seen] == [("ocean", 0), ("ocean", 50), ("ocean", 100), ("ocean", 108)]
    assert seen[-1][2] == 1e30
    # the first check at or after 100 steps is at step 108; the state saved is that step's
    assert _read_p(os.path.join(tmp_path, "ocean_qg_invalid.nc")) == 108
    assert _read_p(os.path.join(tmp_path, "ocean_qg_restart.nc")) == 72
    assert not os.path.exists(os.path.join(tmp_path, "ocean_qg_lastday.nc"))


def test_monitor_cadence(tmp_path):
    clk = Clock(dta=1200.0, nstr=1, tini=0.0, trun=2.0 / 365.0)
    o = FakeQG("ocn")
    seen = []
    run(clk, str(tmp_path), ocn=Basin("ocean", o, nocmon=24), monitor=lambda name, nt, d: seen.append((name, nt)))
    assert seen == [("ocean", nt) for nt in range(0, 145, 24)]
