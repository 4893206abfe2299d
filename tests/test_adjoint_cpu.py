"""CPU checks of the host parts of the gradient path: the wind-stress adjoint,
the kinetic-energy objective's gradient, and the checkpointing schedules."""
from functools import lru_cache

import numpy as np

from openqg_amd.adjoint import Trajectory, _rev_cost, _revolve_split, ekei1
from openqg_amd.forcing import ekman_wekp, ekman_wekp_adjoint


def test_ekman_wekp_adjoint_is_the_transpose():
    rng = np.random.default_rng(0)
    ny, nx, dx, dy, f0 = 13, 17, 3.0e4, 2.5e4, 1.1e-4
    tx, ty = rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx))
    lw = rng.standard_normal((ny, nx))
    w, _, _ = ekman_wekp(nx, ny, dx, dy, f0, False, tx, ty)
    ax, ay = ekman_wekp_adjoint(nx, ny, dx, dy, f0, lw)
    lhs = float(np.sum(lw * w))
    rhs = float(np.sum(ax * tx) + np.sum(ay * ty))
    assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0)


def test_ekei1_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = rng.standard_normal((11, 9))
    args = (3.0e4, 3.0e4, 1.0e-4, 0.5, 1000.0, 300.0)
    J, g = ekei1(p, *args)
    for _ in range(4):
        d = rng.standard_normal(p.shape)
        eps = 1e-6
        fd = (ekei1(p + eps * d, *args)[0] - ekei1(p - eps * d, *args)[0]) / (2 * eps)
        assert abs(fd - np.sum(g * d)) <= 1e-7 * abs(fd)
    # J is a quadratic form: J(p) = 0.5 <grad(p), p>
    assert abs(J - 0.5 * np.sum(g * p)) <= 1e-12 * J


class _FakeAdj:
    """Stands in for AdjointState: the 'state' is the index of the step it
    follows.  Records every snapshot and which step each adjoint step
    reverses (the loaded state must be the one before that step)."""

    def __init__(self):
        self.rec = {"state": 0}
        self.slots = {}
        self.reversed = []
        self.lam_at = None

    def alloc(self, n):
        self.nslots = n

    def save(self, slot):
        assert slot < self.nslots
        self.slots[slot] = self.rec["state"]

    def load(self, slot):
        self.rec["state"] = self.slots[slot]

    def advance(self, tdt, nsteps, step0, supp):
        assert self.rec["state"] == step0 and nsteps > 0
        self.rec["state"] += nsteps

    def record(self, tdt, nsteps, step0, supp, slot0):
        assert self.rec["state"] == step0
        for i in range(nsteps):
            self.save(slot0 + i)
            self.rec["state"] += 1

    def step(self, tdt, suppressed):
        nt = self.rec["state"] + 1            # the step whose input is loaded
        assert nt == self.lam_at, (nt, self.lam_at)
        assert suppressed == ((nt - 1) % 5 == 0)
        self.reversed.append(nt)
        self.lam_at -= 1

    def sweep(self, tdt, nsteps, step0, supp, slot0):
        for i in range(nsteps - 1, -1, -1):
            self.load(slot0 + i)
            self.step(tdt, (step0 + i) % supp == 0)
        self.load(slot0)


def _sweep(n, snapshots):
    a = _FakeAdj()
    t = Trajectory(a, 1.0, n, 5, snapshots)
    t.forward()
    assert a.rec["state"] == n
    a.lam_at = n
    t.reverse()
    assert a.reversed == list(range(n, 0, -1))
    assert a.rec["state"] == 0
    return t.forward_steps


def test_checkpointing_reverses_every_step_once_in_order():
    assert _sweep(17, None) == 17
    for n, s in ((1, 1), (2, 1), (17, 1), (17, 3), (64, 4), (300, 8), (2160, 32)):
        steps = _sweep(n, s)
        assert steps >= n
        # the schedule's cost plus the forward sweep's last stretch to the end
        assert steps <= _rev_cost(n, s) + n
    # 32 free slots besides the initial state's (33 checkpoints) for 2160
    # steps: repetition number 3, Griewank's minimum t = 3n - C(36, 2)
    assert _rev_cost(2160, 32) == 3 * 2160 - 630
    assert _sweep(2160, 32) <= 5850 + 8


@lru_cache(maxsize=None)
def _dp(n, s):
    """Brute-force minimum forward steps of Trajectory._rev's recursion."""
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    return min(m + _dp(n - m, s - 1) + _dp(m, s) for m in range(1, n))


def test_revolve_cost_closed_form_and_split_are_optimal():
    for s in range(0, 5):
        for n in range(1, 61):
            assert _rev_cost(n, s) == _dp(n, s), (n, s)
    for s in range(1, 5):
        for n in range(2, 61):
            m = _revolve_split(n, s)
            assert 1 <= m <= n - 1
            assert m + _dp(n - m, s - 1) + _dp(m, s) == _dp(n, s), (n, s, m)


def test_library_revolve_schedule_matches():
    """The library's schedule (oqg_revolve_cost / oqg_revolve_split, host only)
    against the brute force and the Python schedule."""
    import openqg_amd

    lib = openqg_amd.load()
    for s in range(0, 5):
        for n in range(1, 61):
            assert lib.oqg_revolve_cost(n, s) == _dp(n, s), (n, s)
    for s in range(1, 5):
        for n in range(2, 61):
            m = lib.oqg_revolve_split(n, s)
            assert m + _dp(n - m, s - 1) + _dp(m, s) == _dp(n, s), (n, s, m)
    for n, s in ((2160, 32), (700, 8), (300, 3)):
        assert lib.oqg_revolve_cost(n, s) == _rev_cost(n, s)
        assert lib.oqg_revolve_split(n, s) == _revolve_split(n, s)
    assert lib.oqg_revolve_split(1, 3) == -1 and lib.oqg_revolve_cost(-1, 0) == -1
