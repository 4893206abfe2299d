"""The run loop: clock, step cadence, validity checks and restart dumps.

Mirrors `run` (src/openqg.F90:452-512) with `dynamic_step` (src/driver.F90:36-87)
and the validity part of `diagnostic_step` (src/driver.F90:191-264) for an
ocean basin, an atmosphere basin or both, each a `QG` whose forcing is
resident on the device (`QG.set_forcing` / `QG.set_wind`).  Without a coupler
the two basins do not interact and each advances on its own stream of steps:

- atmosphere: one step of tdta = 2 dta every atmosphere step nt, then
  suppress_comp_mode when mod(nt-1, 100) == 0;
- ocean: one step of tdto = 2 nstr dta when mod(nt-1, nstr) == 0, then
  suppress_comp_mode when mod(nt-1, 25 nstr) == 0.

Between two steps at which something happens on the host (a validity check,
a monitor sample, a restart dump, the end of the run) every basin replays its
steps as cached HIP graphs (`QG.run`), so the host touches the state only at
those steps.  At the other steps the result equals stepping one at a time.

With a coupler (`openqg_amd.coupled.Coupler`) the same cadence runs through
`oqg_cpl_run`: both mixed layers and the exchanges of `dynamic_step` join the
steps, replayed from one captured graph between host events.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from .state import save_aml, save_oml, save_qg

SECDAY = 86400.0            # src/util/constants.F90:17-19
DAYSYR = 365.0
SECSYR = SECDAY * DAYSYR
OCN_SUPP = 25               # ocean suppression every 25 ocean steps (openqg.F90:486)
ATM_SUPP = 100              # atmosphere suppression every 100 steps (openqg.F90:487)


def nint(x: float) -> int:
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


@dataclass
class Clock:
    """clock_type (src/types/clock.F90:10-40) and its derived values
    (load_clock, :45-95).  dta in seconds, tini and trun in years, resday
    (restart dump interval, days; <= 0: none before the end) as in the clock
    file."""

    dta: float
    nstr: int
    tini: float
    trun: float
    resday: float = 0.0

    def __post_init__(self):
        if self.nstr < 1 or not self.dta > 0:
            raise ValueError("need dta > 0 and nstr >= 1")
        self.dto = self.nstr * self.dta
        self.tdto = 2.0 * self.dto
        self.tdta = 2.0 * self.dta
        tend = self.tini + self.trun
        self.nsteps0 = nint(self.tini * SECSYR / self.dta)
        self.nsteps = nint(tend * SECSYR / self.dta)
        self.ntsrun = self.nsteps - self.nsteps0
        self.noutre = self.days_to_steps(self.resday) if self.resday > 0.0 else self.ntsrun + 1

    def days_to_steps(self, days: float) -> int:
        return nint(days * SECDAY / self.dta)   # clock.F90:97-104


@dataclass
class Valids:
    """valids_type's check (src/diagnostics/valsubs.F90:31-37,46-139): every
    nvalid atmosphere steps, the run stops when abs(min) or abs(max) of p
    reaches pext or of q reaches qext.  Non-finite values also stop it."""

    nvalid: int
    pext: float
    qext: float

    def ok(self, d: dict) -> bool:
        if d["nonfinite"] > 0:
            return False
        pmin, pmax = float(np.min(d["pmin"])), float(np.max(d["pmax"]))
        qmin, qmax = float(np.min(d["qmin"])), float(np.max(d["qmax"]))
        return not (abs(pmin) >= self.pext or abs(pmax) >= self.pext or
                    abs(qmin) >= self.qext or abs(qmax) >= self.qext)


class InvalidSolution(RuntimeError):
    """A validity check failed (the reference prints and stops, openqg.F90:495-501)."""


def ocean_steps(nt0: int, nt1: int, nstr: int):
    """(first ocean step index, count) of the ocean steps taken at atmosphere
    steps nt0+1 .. nt1: the ocean steps at nt with mod(nt-1, nstr) == 0, the
    ocean step index of such an nt being (nt-1)/nstr."""
    first = (nt0 + nstr - 1) // nstr
    last = (nt1 - 1) // nstr
    return first, max(0, last - first + 1)


@dataclass
class Basin:
    """A QG basin in the run loop: its name in file names, its `QG` and its
    optional validity check and monitor cadence (steps)."""

    name: str
    qg: object
    valids: Optional[Valids] = None
    nocmon: int = 0


def run(clk: Clock, outdir: str, ocn: Optional[Basin] = None, atm: Optional[Basin] = None,
        monitor: Optional[Callable[[str, int, dict], None]] = None, coupler=None) -> int:
    """The run loop of src/openqg.F90:479-512.

    Writes `<name>_qg_restart.nc` every clk.noutre steps and
    `<name>_qg_lastday.nc` at the end (src/openqg.F90:503-512).  When a
    validity check fails, writes `<name>_qg_invalid.nc` for each basin and
    raises InvalidSolution.  `monitor(name, nt, diag)` receives `QG.diagnose`
    every basin.nocmon steps (qg_monitor_step's cadence).  Returns the last
    step number.

    With `coupler` (an openqg_amd.coupled.Coupler of ocn and atm) the two basins
    advance together through dynamic_step (src/driver.F90:36-87) with both
    mixed layers, and the restart files include the mixed layers'
    `<basin>_ml_*.nc` (save_all, src/openqg.F90:518-543)."""
    basins = [b for b in (atm, ocn) if b is not None]
    if coupler is not None and (ocn is None or atm is None):
        raise ValueError("a coupled run needs both basins")

    def save_all(suffix):
        for b in basins:
            save_qg(outdir, f"{b.name}_qg_{suffix}", b.qg)
        if coupler is not None:
            save_oml(outdir, f"{ocn.name}_ml_{suffix}", coupler.oml)
            save_aml(outdir, f"{atm.name}_ml_{suffix}", coupler)
    dts = {}
    if atm is not None:
        dts[id(atm)] = clk.dta
    if ocn is not None:
        dts[id(ocn)] = clk.dto

    def cadences():
        for b in basins:
            if b.valids is not None:
                yield b.valids.nvalid
            if monitor is not None and b.nocmon > 0:
                yield b.nocmon
        yield clk.noutre

    last = {}   # per basin: (tsec, et2m) of its previous monitor sample

    def sample(b, nt, d):
        """A monitor sample: the state diagnostics, the rest of qg_diagno's
        energetics (QG.diagnose_energy) and pket from the previous sample's
        et2m (qg_monitor.F90:266-272, tsec = nt dta)."""
        d = dict(d)
        if hasattr(b.qg, "diagnose_energy"):
            d.update(b.qg.diagnose_energy())
        if "et2m" in d:
            tsec = nt * clk.dta
            et2 = np.asarray(d["et2m"])
            prev = last.get(id(b))
            if prev is not None and tsec > prev[0]:
                d["pket"] = 0.5 * b.qg.rho * np.asarray(b.qg.gp) * (et2 - prev[1]) / (tsec - prev[0])
            else:
                d["pket"] = np.zeros_like(et2)
            last[id(b)] = (tsec, et2)
        return d

    def diagnostic_step(nt):
        """driver.F90:191-264: valids for the atmosphere, then the ocean; a
        failed check forces every active monitor's sample at this step
        (force = .not. solnok, qg_monitor_step: qg_monitor.F90:134-141); the
        qg monitors for the ocean, then the atmosphere; then the invalid
        dump."""
        ntdone = nt - clk.nsteps0
        solnok = True
        diags = {}
        for b in basins:
            if b.valids is not None and ntdone % b.valids.nvalid == 0:
                d = diags[id(b)] = b.qg.diagnose(dts[id(b)])
                if not b.valids.ok(d):
                    solnok = False
        force = not solnok
        if monitor is not None:
            for b in (ocn, atm):
                if b is not None and b.nocmon > 0 and (force or ntdone % b.nocmon == 0):
                    d = diags.get(id(b))
                    monitor(b.name, nt, sample(b, nt, d if d is not None else b.qg.diagnose(dts[id(b)])))
        if not solnok:
            save_all("invalid.nc")
            raise InvalidSolution(f"invalid values at nt = {nt}, day {nt * clk.dta / SECDAY:.2f}")

    steps = sorted(set(c for c in cadences() if c > 0))
    nt = clk.nsteps0
    diagnostic_step(nt)
    while nt < clk.nsteps:
        ntdone = nt - clk.nsteps0
        nxt = min([clk.ntsrun] + [(ntdone // c + 1) * c for c in steps])
        nt1 = clk.nsteps0 + nxt
        if coupler is not None:
            coupler.run(nt1 - nt, nt, clk.tdto, clk.tdta, clk.nstr, OCN_SUPP * clk.nstr, ATM_SUPP)
        else:
            if atm is not None:
                atm.qg.run(clk.tdta, nt1 - nt, nt, ATM_SUPP)
            if ocn is not None:
                first, count = ocean_steps(nt, nt1, clk.nstr)
                if count:
                    ocn.qg.run(clk.tdto, count, first, OCN_SUPP)
        nt = nt1
        diagnostic_step(nt)
        if (nt - clk.nsteps0) % clk.noutre == 0:
            save_all("restart.nc")
    save_all("lastday.nc")
    return nt
