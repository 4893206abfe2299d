"""Gradients of the QG model through the hand-written adjoint of step_qg
(box or zonally periodic channel).

The objective is the layer-1 kinetic energy of qg_monitor (ekei(1),
src/diagnostics/qg_monitor.F90:297-302,331-333) after N steps:
    J = 0.5 rho H1 int_P_dA(u^2 + v^2),  u = -dPdy_2bc(p1)/f0, v = dPdx_2bc(p1)/f0
(numerics.F90:470-513 for the boundary differences, :165-204 for int_P_dA).
Controls: the kinematic wind stress (taux, tauy) at p points, through
ekman_from_tau + map_T_to_P (on the device, oqg_set_wind / oqg_adj_wind), and
the interior PV of the initial state, whose pressure is one inversion of it
(dpi = 0), with qm = q and pm = p.

The forward run keeps device snapshots of the state; the reverse sweep applies
oqg_adj_step from them.  With `snapshots=None` every state is kept (HBM holds
thousands of them; oqg_adj_record / oqg_adj_sweep run both sweeps from C);
with a number, binomial checkpointing (Griewank's revolve schedule)
recomputes the forward from the stored ones with graph-replayed oqg_run.

In a channel the fields are periodic in x with column nxp-1 a copy of column
0: the controls given keep that copy, and the gradients returned are those of
the periodic fields (the derivative with respect to a periodic point sits in
its column 0..nxp-2, column nxp-1 holds 0); the stress gradient includes the
paths through the boundary stress integrals txis, txin.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache
from math import comb
from typing import Optional

import numpy as np

from . import _lib
from ._lib import OQG_HOST, check
from .qg import QG


# ------------------------------------------------------------------ objective

def _trap_weights(nyp: int, nxp: int, dx: float, dy: float) -> np.ndarray:
    w = np.ones((nyp, nxp))
    w[0, :] *= 0.5
    w[-1, :] *= 0.5
    w[:, 0] *= 0.5
    w[:, -1] *= 0.5
    return w * dx * dy


def _ddx(p, dx, alpha):
    d = np.zeros_like(p)
    d[:, 1:-1] = (p[:, 2:] - p[:, :-2]) / (2.0 * dx)
    d[:, 0] = (p[:, 1] - p[:, 0]) / (dx * (0.5 * alpha + 1.0))
    d[:, -1] = (p[:, -1] - p[:, -2]) / (dx * (0.5 * alpha + 1.0))
    d[0, :] = 0.0
    d[-1, :] = 0.0
    return d


def _ddy(p, dy, alpha):
    d = np.zeros_like(p)
    d[1:-1, :] = (p[2:, :] - p[:-2, :]) / (2.0 * dy)
    d[0, :] = (p[1, :] - p[0, :]) / (dy * (0.5 * alpha + 1.0))
    d[-1, :] = (p[-1, :] - p[-2, :]) / (dy * (0.5 * alpha + 1.0))
    d[:, 0] = 0.0
    d[:, -1] = 0.0
    return d


def _ddx_T(g, dx, alpha):
    """Transpose of _ddx (box)."""
    g = g.copy()
    g[0, :] = 0.0
    g[-1, :] = 0.0
    out = np.zeros_like(g)
    out[:, 2:] += g[:, 1:-1] / (2.0 * dx)
    out[:, :-2] -= g[:, 1:-1] / (2.0 * dx)
    c = dx * (0.5 * alpha + 1.0)
    out[:, 1] += g[:, 0] / c
    out[:, 0] -= g[:, 0] / c
    out[:, -1] += g[:, -1] / c
    out[:, -2] -= g[:, -1] / c
    return out


def _ddy_T(g, dy, alpha):
    """Transpose of _ddy (box)."""
    g = g.copy()
    g[:, 0] = 0.0
    g[:, -1] = 0.0
    out = np.zeros_like(g)
    out[2:, :] += g[1:-1, :] / (2.0 * dy)
    out[:-2, :] -= g[1:-1, :] / (2.0 * dy)
    c = dy * (0.5 * alpha + 1.0)
    out[1, :] += g[0, :] / c
    out[0, :] -= g[0, :] / c
    out[-1, :] += g[-1, :] / c
    out[-2, :] -= g[-1, :] / c
    return out


def ekei1(p1: np.ndarray, dx: float, dy: float, fnot: float, bcco: float, rho: float, h1: float):
    """Layer-1 kinetic energy (qg_monitor ekei(1)) of the pressure p1 (nyp, nxp)
    and its gradient with respect to p1."""
    w = _trap_weights(*p1.shape, dx, dy)
    u = -_ddy(p1, dy, bcco) / fnot
    v = _ddx(p1, dx, bcco) / fnot
    J = 0.5 * rho * h1 * float(np.sum(w * (u * u + v * v)))
    gu = rho * h1 * w * u
    gv = rho * h1 * w * v
    grad = -_ddy_T(gu, dy, bcco) / fnot + _ddx_T(gv, dx, bcco) / fnot
    return J, grad


# ------------------------------------------------------------------ device adjoint state

class AdjointState:
    """The adjoint buffers and state snapshots of one basin."""

    def __init__(self, qg: QG):
        self.qg = qg
        self._lib = qg._lib
        check(self._lib.oqg_adj_init(qg._h), qg._h, "oqg_adj_init")
        self.nslots = 0

    def _c(self, rc, what):
        check(rc, self.qg._h, what)

    def alloc(self, nslots: int):
        if nslots > self.nslots:
            self._c(self._lib.oqg_snap_alloc(self.qg._h, int(nslots)), "oqg_snap_alloc")
            self.nslots = nslots

    def save(self, slot: int):
        self._c(self._lib.oqg_snap_save(self.qg._h, int(slot)), "oqg_snap_save")

    def load(self, slot: int):
        self._c(self._lib.oqg_snap_load(self.qg._h, int(slot)), "oqg_snap_load")

    def set(self, lq=None, lqm=None, lp=None, lpm=None):
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (lq, lqm, lp, lpm)]
        ptrs = [None if a is None else a.ctypes.data_as(ctypes.c_void_p) for a in keep]
        self._c(self._lib.oqg_adj_set(self.qg._h, *ptrs, OQG_HOST), "oqg_adj_set")

    def set_constr(self, v):
        """The adjoints of the constraint scalars, in QG.get_constr's packing."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._c(self._lib.oqg_adj_set_constr(self.qg._h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                "oqg_adj_set_constr")

    def get_constr(self) -> np.ndarray:
        out = np.zeros(self._lib.oqg_constr_len(self.qg._h))
        self._c(self._lib.oqg_adj_get_constr(self.qg._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                "oqg_adj_get_constr")
        return out

    def txi(self) -> np.ndarray:
        """The accumulated adjoints of the channel's txis, txin."""
        out = np.zeros(2)
        self._c(self._lib.oqg_adj_get_txi(self.qg._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                "oqg_adj_get_txi")
        return out

    def seed_ke(self, rho: float) -> float:
        """J = ekei(1) of the loaded state and the adjoint seeded with dJ/dp
        (oqg_adj_seed_ke, on the device)."""
        J = ctypes.c_double()
        self._c(self._lib.oqg_adj_seed_ke(self.qg._h, float(rho), ctypes.byref(J)), "oqg_adj_seed_ke")
        return J.value

    def get(self):
        out = [np.empty(self.qg.shape) for _ in range(4)]
        self._c(self._lib.oqg_adj_get(self.qg._h, *(o.ctypes.data_as(ctypes.c_void_p) for o in out), OQG_HOST),
                "oqg_adj_get")
        return tuple(out)

    def get_lq(self) -> np.ndarray:
        """The adjoint of q alone."""
        out = np.empty(self.qg.shape)
        self._c(self._lib.oqg_adj_get(self.qg._h, out.ctypes.data_as(ctypes.c_void_p), None, None, None, OQG_HOST),
                "oqg_adj_get")
        return out

    def step(self, tdt: float, suppressed: bool):
        self._c(self._lib.oqg_adj_step(self.qg._h, float(tdt), int(bool(suppressed))), "oqg_adj_step")

    def advance(self, tdt: float, nsteps: int, step0: int, supp_every: int):
        """Forward steps step0+1 .. step0+nsteps without snapshots (oqg_run)."""
        self.qg.run(tdt, nsteps, step0, supp_every)

    def record(self, tdt: float, nsteps: int, step0: int, supp_every: int, slot0: int):
        """Forward steps saving the state before each in slots slot0 + i."""
        self._c(self._lib.oqg_adj_record(self.qg._h, float(tdt), int(nsteps), int(step0), int(supp_every),
                                         int(slot0)), "oqg_adj_record")

    def sweep(self, tdt: float, nsteps: int, step0: int, supp_every: int, slot0: int):
        """Adjoint steps back through a record; leaves slot0's state loaded."""
        self._c(self._lib.oqg_adj_sweep(self.qg._h, float(tdt), int(nsteps), int(step0), int(supp_every),
                                        int(slot0)), "oqg_adj_sweep")

    def revolve_forward(self, tdt: float, nsteps: int, supp_every: int, free_slots: int):
        """Forward sweep of the binomial checkpointing schedule, run by the library."""
        self._c(self._lib.oqg_adj_revolve_forward(self.qg._h, float(tdt), int(nsteps), int(supp_every),
                                                  int(free_slots)), "oqg_adj_revolve_forward")

    def revolve_reverse(self) -> int:
        """Its reverse sweep; returns the forward steps of the whole schedule."""
        n = ctypes.c_longlong(0)
        self._c(self._lib.oqg_adj_revolve_reverse(self.qg._h, ctypes.byref(n)), "oqg_adj_revolve_reverse")
        return int(n.value)

    def wekp(self) -> np.ndarray:
        out = np.empty((self.qg.b.nyp, self.qg.b.nxp))
        self._c(self._lib.oqg_adj_get_wekp(self.qg._h, out.ctypes.data_as(ctypes.c_void_p), OQG_HOST),
                "oqg_adj_get_wekp")
        return out

    def zero_wekp(self):
        self._c(self._lib.oqg_adj_zero_wekp(self.qg._h), "oqg_adj_zero_wekp")

    def wind(self):
        """The adjoints of (taux, tauy) from the accumulated adjoint of wekp,
        through the transpose of ekman_from_tau on the device."""
        out = [np.empty((self.qg.b.nyp, self.qg.b.nxp)) for _ in range(2)]
        self._c(self._lib.oqg_adj_wind(self.qg._h, *(o.ctypes.data_as(ctypes.c_void_p) for o in out), OQG_HOST),
                "oqg_adj_wind")
        return out[0], out[1]

    def invert(self):
        self._c(self._lib.oqg_invert(self.qg._h), "oqg_invert")

    def adj_invert(self):
        self._c(self._lib.oqg_adj_invert(self.qg._h), "oqg_adj_invert")


# ------------------------------------------------------------------ checkpointing

def _beta(c: int, r: int) -> int:
    return comb(c + r, c) if r >= 0 else 0


def _rev_cost(n: int, s: int) -> int:
    """Forward steps that reversing n steps costs from a stored state with s
    further free slots (Trajectory._rev): Griewank's t(n, c) = r n - beta(c+1,
    r-1) with c = s + 1 checkpoints and r the least repetition number with
    beta(c, r) = C(c + r, c) >= n (Griewank and Walther, ACM TOMS 26, 2000)."""
    if n <= 1:
        return 0
    c, r = s + 1, 0
    while _beta(c, r) < n:
        r += 1
    return r * n - _beta(c + 1, r - 1)


def _cost_vec(x: np.ndarray, c: int) -> np.ndarray:
    """_rev_cost over an array of step counts x (>= 1) for c = s + 1 checkpoints."""
    nmax = int(x.max())
    betas = [1]
    while betas[-1] < nmax:
        betas.append(_beta(c, len(betas)))
    r = np.searchsorted(np.array(betas, dtype=np.float64), x, side="left")
    prev = np.array([0] + [_beta(c + 1, k) for k in range(len(betas) - 1)], dtype=np.float64)
    return np.where(x <= 1, 0.0, r * x - prev[r])


@lru_cache(maxsize=None)
def _revolve_split(n: int, s: int) -> int:
    """The advance m (1 <= m < n) before the next checkpoint when reversing n
    steps with s >= 1 free slots: the m that minimises
    m + cost(n - m, s - 1) + cost(m, s)."""
    if n <= 2:
        return 1
    m = np.arange(1, n, dtype=np.float64)
    total = m + _cost_vec(n - m, s) + _cost_vec(m, s + 1)
    return int(m[int(np.argmin(total))])


class Trajectory:
    """Forward run of `nsteps` from the loaded state with the schedule of
    step_qg + suppress_comp_mode, and its reverse sweep."""

    def __init__(self, adj: AdjointState, tdt: float, nsteps: int, supp_every: int, snapshots: Optional[int]):
        self.adj, self.tdt, self.n, self.supp = adj, float(tdt), int(nsteps), int(supp_every)
        self.snapshots = snapshots
        self.forward_steps = 0   # counts recomputation too

    def _suppressed(self, nt: int) -> bool:
        return self.supp > 0 and (nt - 1) % self.supp == 0

    def forward(self):
        """Run forward; returns after step n with the final state loaded.  With
        checkpointing, the sweep places the checkpoints of the schedule's
        rightmost path on the way (the first advance of every level)."""
        a = self.adj
        if self.snapshots is None:
            a.alloc(max(self.n, 1))
            a.record(self.tdt, self.n, 0, self.supp, 0)
            self.forward_steps += self.n
            return
        a.alloc(self.snapshots + 1)
        if hasattr(a, "revolve_forward"):
            # the library runs the same schedule from C, with graph-replayed steps
            a.revolve_forward(self.tdt, self.n, self.supp, self.snapshots)
            return
        a.save(0)
        self._left = []
        lo, slot, free = 0, 0, self.snapshots
        while self.n - lo > 1 and free > 0:
            m = lo + _revolve_split(self.n - lo, free)
            self._advance(lo, m)
            a.save(slot + 1)
            self._left.append((lo, m, slot, free))
            lo, slot, free = m, slot + 1, free - 1
        self._advance(lo, self.n)
        self._tail = (lo, slot, free)

    def reverse(self):
        """With the adjoint of the final state set, sweep back to the initial state."""
        a = self.adj
        if self.snapshots is None:
            a.sweep(self.tdt, self.n, 0, self.supp, 0)
            return
        if hasattr(a, "revolve_reverse"):
            self.forward_steps = a.revolve_reverse()
            return
        lo, slot, free = self._tail
        self._rev(lo, self.n, slot, free)
        for lo, m, slot, free in reversed(self._left):
            self._rev(lo, m, slot, free)
        a.load(0)

    def _advance(self, lo: int, hi: int):
        if hi > lo:
            self.adj.advance(self.tdt, hi - lo, lo, self.supp)
            self.forward_steps += hi - lo

    def _rev(self, lo: int, hi: int, slot: int, free: int):
        """State lo is in `slot`; the adjoint holds state hi; leave it at state lo."""
        a = self.adj
        n = hi - lo
        if n == 1:
            a.load(slot)
            a.step(self.tdt, self._suppressed(hi))
            return
        if free == 0:
            for k in range(hi - 1, lo - 1, -1):
                a.load(slot)
                self._advance(lo, k)
                a.step(self.tdt, self._suppressed(k + 1))
            return
        m = lo + _revolve_split(n, free)
        a.load(slot)
        self._advance(lo, m)
        a.save(slot + 1)
        self._rev(m, hi, slot + 1, free - 1)
        self._rev(lo, m, slot, free)


def gradient(qg: QG, q0: np.ndarray, taux: np.ndarray, tauy: np.ndarray, tdt: float, nsteps: int,
             supp_every: int = 25, snapshots: Optional[int] = None, rho: float = 1000.0):
    """J = ekei(1) after nsteps from the state whose interior PV is q0
    (nl, nyp, nxp; ring ignored) under the kinematic wind stress (taux, tauy);
    returns (J, dJ/dtaux, dJ/dtauy, dJ/dq0 (interior, ring 0), forward steps run)."""
    b = qg.b
    adj = AdjointState(qg)
    ent = np.zeros((max(b.nl - 1, 1), b.nyp, b.nxp))
    qg.set_forcing(ent, None, 0.0, 0.0)
    qg.set_wind(taux, tauy)   # wekp on the device
    # the inversion writes p from q alone and then sets qm = q, pm = p
    qg.set_state(q=q0)
    qg.set_constr(np.zeros(qg._lib.oqg_constr_len(qg._h)))
    adj.invert()
    traj = Trajectory(adj, tdt, nsteps, supp_every, snapshots)
    traj.forward()
    J = adj.seed_ke(rho)
    adj.zero_wekp()
    traj.reverse()
    gtx, gty = adj.wind()     # the transpose of the wind forcing on the device
    adj.adj_invert()
    gq0 = adj.get_lq()
    return J, gtx, gty, gq0, traj.forward_steps


def gradient_batch(qgs, q0s, tauxs, tauys, tdt: float, nsteps: int, supp_every: int = 25,
                   snapshots: Optional[int] = None, rho: float = 1000.0, concurrent: bool = True):
    """`gradient` for a batch of members, member i on its own context qgs[i]
    (each with its own stream) from (q0s[i], tauxs[i], tauys[i]).

    With `concurrent` every member runs in its own host thread: the library
    releases the interpreter lock inside each call and a context touches only
    its own stream and memory, so the members' kernels overlap on the device.
    The results equal those of running the members one after another.
    Returns the list of `gradient`'s tuples in member order."""
    n = len(qgs)
    if not (len(q0s) == len(tauxs) == len(tauys) == n):
        raise ValueError("one q0, taux and tauy per member")
    if len({id(q) for q in qgs}) != n:
        raise ValueError("every member needs its own context")
    run1 = lambda i: gradient(qgs[i], q0s[i], tauxs[i], tauys[i], tdt, nsteps, supp_every, snapshots, rho)
    if not concurrent or n == 1:
        return [run1(i) for i in range(n)]
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=n) as ex:
        return list(ex.map(run1, range(n)))


def objective(qg: QG, q0, taux, tauy, tdt: float, nsteps: int, supp_every: int = 25, rho: float = 1000.0) -> float:
    """The forward map of `gradient` alone (for finite-difference checks)."""
    b = qg.b
    ent = np.zeros((max(b.nl - 1, 1), b.nyp, b.nxp))
    qg.set_forcing(ent, None, 0.0, 0.0)
    qg.set_wind(taux, tauy)
    qg.set_state(q=q0)
    qg.set_constr(np.zeros(qg._lib.oqg_constr_len(qg._h)))
    check(qg._lib.oqg_invert(qg._h), qg._h, "oqg_invert")
    qg.run(tdt, nsteps, 0, supp_every)
    d = _lib.OqgDiag()
    check(qg._lib.oqg_diagnose(qg._h, 0.5 * float(tdt), float(rho), ctypes.byref(d)), qg._h, "oqg_diagnose")
    return d.ekei[0]   # the same device sums as the gradient's objective


def make_autograd(qg: QG, tdt: float, nsteps: int, supp_every: int = 25, snapshots: Optional[int] = None):
    """A torch.autograd.Function J(taux, tauy, q0) backed by `gradient`."""
    import torch

    class _KE(torch.autograd.Function):
        @staticmethod
        def forward(ctx, taux, tauy, q0):
            J, gx, gy, gq, _ = gradient(qg, q0.detach().cpu().numpy(), taux.detach().cpu().numpy(),
                                        tauy.detach().cpu().numpy(), tdt, nsteps, supp_every, snapshots)
            ctx.save_for_backward(*(torch.from_numpy(a).to(taux.device) for a in (gx, gy, gq)))
            return torch.tensor(J, dtype=torch.float64, device=taux.device)

        @staticmethod
        def backward(ctx, gout):
            gx, gy, gq = ctx.saved_tensors
            return gout * gx, gout * gy, gout * gq

    return _KE.apply
