"""Batched ensembles of basins over liboqg.so (oqg_ens_* in include/oqg.h).

The reference runs an ensemble as separate model runs, one basin per step_qg
call (src/driver.F90:107).  Here independent basins of any sizes and layer
counts advance together: every phase of the step is one launch over the
blocks of all members (per class of kernel template parameters), replayed from
HIP graphs.  A member's trajectory is the one it would follow stepped alone.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from ._lib import ENS_LARGEST_FIRST, ENS_ONE_YSOLVE, OqgError, load  # noqa: F401
from .qg import QG


class Ensemble:
    """Members are `QG` basins (already holding their state and forcing); they
    must outlive the ensemble and must not be stepped on their own while it runs."""

    def __init__(self, members: Sequence[QG]):
        self._lib = load()
        self.members = list(members)
        n = len(self.members)
        arr = (ctypes.c_void_p * n)(*[m._h for m in self.members])
        h = ctypes.c_void_p()
        rc = self._lib.oqg_ens_create(arr, n, ctypes.byref(h))
        if rc != 0:
            raise OqgError(f"oqg_ens_create failed (code {rc}): {self._lib.oqg_ens_last_error(None).decode()}")
        self._h = h

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise OqgError(f"{what} failed (code {rc}): {self._lib.oqg_ens_last_error(self._h).decode()}")

    @staticmethod
    def _arrays(tdt, nsteps, n):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(tdt, dtype=np.float64), (n,)))
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(nsteps, dtype=np.int32), (n,)))
        return t, s

    def run(self, tdt, nsteps, step0: int = 0, supp_every: int = 25):
        """Advance member i by nsteps[i] leapfrog steps of 2*dt = tdt[i] (scalars broadcast)."""
        t, s = self._arrays(tdt, nsteps, len(self.members))
        self._check(self._lib.oqg_ens_run(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          s.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(step0),
                                          int(supp_every)), "oqg_ens_run")

    def sync(self):
        self._check(self._lib.oqg_ens_sync(self._h), "oqg_ens_sync")

    @property
    def schedule(self) -> int:
        return int(self._lib.oqg_ens_get_schedule(self._h))

    @schedule.setter
    def schedule(self, value: int):
        self._check(self._lib.oqg_ens_set_schedule(self._h, int(value)), "oqg_ens_set_schedule")

    TILE_OPTIONS = (0, 4, 8, 16)   # 0: each member's default rule

    def tune(self, tdt, trial_steps: int = 8, supp_every: int = 25):
        """Time every launch schedule with every vorticity tile height on the
        members (their state is restored), keep the fastest pair; returns
        (schedule, [ms] indexed t*4 + schedule) and sets `tile_option` to the
        height chosen (TILE_OPTIONS[t])."""
        t, _ = self._arrays(tdt, 0, len(self.members))
        ms = (ctypes.c_double * 16)()
        best = self._lib.oqg_ens_tune(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(trial_steps),
                                      int(supp_every), ms)
        if best < 0:
            self._check(-best, "oqg_ens_tune")
        self.tile_option = self.TILE_OPTIONS[best // 4]
        return int(best % 4), list(ms)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.oqg_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnsembleGroups:
    """Members split into `groups` batched ensembles of about equal work per
    step (largest members first, each to the group with the least work so far),
    each ensemble on its own stream.  `run` enqueues every group before any
    finishes, so one group's long launches (its largest members) overlap the
    other groups' phases instead of leaving the device partly idle at the end
    of each phase.  Members of at least `solo_points` grid points (nxp nyp)
    are not batched: each replays its own graphs (`QG.run`) on its own stream,
    beside the groups; such a member fills the device on its own, and batching
    it only makes the small members' blocks wait behind it.  Each member's
    trajectory is unchanged."""

    def __init__(self, members: Sequence[QG], groups: int = 2, solo_points: Optional[int] = None):
        self.members = list(members)
        n = len(self.members)
        pts = [m.b.nxp * m.b.nyp for m in self.members]
        self.solo = [i for i in range(n) if solo_points is not None and pts[i] >= solo_points]
        rest = [i for i in range(n) if i not in self.solo]
        groups = max(1, min(int(groups), max(len(rest), 1)))
        work = [m.b.nxp * m.b.nyp * m.b.nl for m in self.members]
        load_ = [0] * groups
        idx = [[] for _ in range(groups)]
        for i in sorted(rest, key=lambda i: -work[i]):
            g = int(np.argmin(load_))
            idx[g].append(i)
            load_[g] += work[i]
        self.index = [sorted(g) for g in idx if g]
        self.parts = [Ensemble([self.members[i] for i in g]) for g in self.index]

    def _split(self, v):
        v = np.broadcast_to(np.asarray(v), (len(self.members),))
        return [v[g] for g in self.index]

    def run(self, tdt, nsteps, step0: int = 0, supp_every: int = 25):
        """As Ensemble.run; the solo members and the groups run concurrently."""
        t_all = np.broadcast_to(np.asarray(tdt, dtype=np.float64), (len(self.members),))
        s_all = np.broadcast_to(np.asarray(nsteps), (len(self.members),))
        for i in self.solo:
            if s_all[i] > 0:
                self.members[i].run(float(t_all[i]), int(s_all[i]), step0, supp_every)
        for e, t, s in zip(self.parts, self._split(tdt), self._split(nsteps)):
            e.run(t, s, step0, supp_every)

    def sync(self):
        for i in self.solo:
            self.members[i].sync()
        for e in self.parts:
            e.sync()

    def tune(self, tdt, trial_steps: int = 8, supp_every: int = 25):
        """Ensemble.tune for every group and QG.tune for every solo member;
        returns [(schedule, [ms])] per group."""
        t_all = np.broadcast_to(np.asarray(tdt, dtype=np.float64), (len(self.members),))
        for i in self.solo:
            self.members[i].tune(float(t_all[i]), trial_steps, supp_every)
        return [e.tune(t, trial_steps, supp_every) for e, t in zip(self.parts, self._split(tdt))]

    def close(self):
        for e in self.parts:
            e.close()
