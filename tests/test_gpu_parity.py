"""GPU parity of the HIP step against the CPU oracle (oracle/cpuref.c).

Bar (FP64): relative max-norm error <= 1e-12 on q and p after each step taken
from an identical state, and <= 1e-9 over a 10-day 129^2 trajectory.  The
constraint scalars follow the same bar.  Whole-size (513^2) runs are also
checked through properties that hold at any size (mass constraint identity,
graph replay == eager launches, bitwise determinism).
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle import cpuref as cr
from openqg_amd import workloads as W
from openqg_amd import FLAG_YSOLVE_THOMAS, OqgError

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
TRAJ_TOL = 1e-9


def _setup(w, flags=0, ent_amp=0.0):
    qg = w.make_qg(flags=flags)
    o = cr.Basin(w.params())
    if w.init == "splitmix":
        # q from SplitMix64 deviates, p by one inversion on each side
        w.start(qg)
        o.start_from_q(W.initial_q(w))
    else:
        p = w.initial_p()
        for b in (qg, o):
            b.set_state(p=p, pm=p)
            b.init_state()
    ent, wekp, txis, txin = w.forcing()
    if ent_amp:
        # seeded entrainment on every layer interface (m/s); a channel's
        # duplicate column nxp equals column 1
        ent = ent_amp * np.random.default_rng(w.seed + 100).standard_normal(ent.shape)
        if w.cyclic:
            ent[..., -1] = ent[..., 0]
    qg.set_forcing(ent, wekp, txis, txin)
    return qg, o, (ent, wekp, txis, txin)


def _constr_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b), np.abs(b).max() * 1e-3 + 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def _per_step(w, nsteps, flags=0, tol=STEP_TOL, ent_amp=0.0):
    qg, o, (ent, wekp, txis, txin) = _setup(w, flags, ent_amp)
    print(f"[parity] {w.name} nsteps={nsteps} flags={flags}", flush=True)
    gq, gqm, gp, gpm = qg.get_state()
    oq, oqm, op, opm = o.get_state()
    assert rel_err(gq, oq) <= tol, "init_pv q"
    assert rel_err(gp, op) <= tol, "initial p"
    assert _constr_err(qg.get_constr(), o.get_constr()) <= 1e-10, "init constraints"
    worst = 0.0
    for nt in range(1, nsteps + 1):
        # restart the oracle from the GPU state so each comparison is one step
        gq, gqm, gp, gpm = qg.get_state()
        o.set_state(gq, gqm, gp, gpm)
        o.set_constr(qg.get_constr())
        qg.step(w.tdt)
        o.step(w.tdt, ent, wekp, txis, txin)
        if (nt - 1) % w.supp_every == 0:
            qg.suppress()
            o.suppress()
        gq, gqm, gp, gpm = qg.get_state()
        oq, oqm, op, opm = o.get_state()
        eq, ep = rel_err(gq, oq), rel_err(gp, op)
        ec = _constr_err(qg.get_constr(), o.get_constr())
        worst = max(worst, eq, ep)
        assert eq <= tol, f"step {nt}: q rel err {eq:.3e}"
        assert ep <= tol, f"step {nt}: p rel err {ep:.3e}"
        assert rel_err(gqm, oqm) <= tol and rel_err(gpm, opm) <= tol
        assert ec <= 1e-9, f"step {nt}: constraint rel err {ec:.3e}"
    print(f"[parity] {w.name}: worst per-step rel err {worst:.3e}", flush=True)
    return qg, o


@pytest.mark.parametrize("make", [W.small_box, W.small_channel], ids=["box33", "chan33x17"])
def test_step_parity_small(gpu, make):
    _per_step(make(), 30)


@pytest.mark.parametrize("make", [lambda: W.small_box(n=65), W.small_channel, W.atmos_channel],
                         ids=["box65", "chan33x17", "atmos97x49"])
def test_step_parity_with_entrainment(gpu, make):
    """Entrainment forcing of layers 1 and 2 (vorticity.F90:71-72), the mass
    constraint's ent integral (constraint.F90:167-179) and, in a channel, the
    layer-2 edge rows in the momentum constraints (vorticity.F90:88-93)."""
    _per_step(make(), 12, ent_amp=1e-6)


@pytest.mark.parametrize("nl", [2, 4])
def test_step_parity_layers(gpu, nl):
    _per_step(W.small_box(n=41, nl=nl), 8)
    _per_step(W.small_channel(nx=49, ny=21, nl=nl), 8)


def test_step_parity_mixed_radix(gpu):
    """Transform lengths with both 3 and 5 as factors (120 = 8*3*5 in the box,
    60 = 4*3*5 in the channel): both root blocks of a plan's twiddle table."""
    _per_step(W.small_box(n=61), 6)
    _per_step(W.small_channel(nx=61, ny=31), 6)


@pytest.mark.parametrize("kind,n", [("box", 57), ("box", 45), ("box", 40), ("chan", 53), ("chan", 57),
                                    ("chan", 92)],
                         ids=["box57_L112=16x7", "box45_L88=8x11", "box40_L78=2x3x13", "chan53_L52=4x13",
                              "chan57_L56=8x7", "chan92_L91=7x13"])
def test_step_parity_prime_lengths(gpu, kind, n):
    """Transform lengths with a factor 7, 11 or 13 run the LDS block kernels
    with those radices (FFTW's any-length transforms, inhomog.F90:103-122)."""
    w = W.small_box(n=n) if kind == "box" else W.small_channel(nx=n, ny=25)
    _per_step(w, 6)


def test_prime_lengths_rejected_by_batches_and_adjoint(gpu):
    """The batched ensemble and the adjoint take lengths of 2, 3 and 5 only:
    both refuse a basin with a radix-7 plan with an error, not a wrong result."""
    from openqg_amd.adjoint import AdjointState
    from openqg_amd.ensemble import Ensemble

    w = W.small_box(n=57)
    qg = w.make_qg()
    with pytest.raises(OqgError, match="factor into 2, 3 and 5"):
        AdjointState(qg)
    with pytest.raises(OqgError, match="factor into 2, 3 and 5"):
        Ensemble([qg, W.small_box(n=33).make_qg()])


def test_step_parity_rectangular_box(gpu):
    w = W.small_box(n=33)
    w.nxp, w.nyp, w.name = 65, 97, "box65x97"
    _per_step(w, 6)


def test_step_parity_thomas_variant(gpu):
    _per_step(W.small_box(n=65), 6, flags=FLAG_YSOLVE_THOMAS)
    _per_step(W.small_channel(), 6, flags=FLAG_YSOLVE_THOMAS)


@pytest.mark.parametrize("make", [W.c1_box129, W.atmos_channel], ids=["box129", "atmos97x49"])
def test_step_parity_configs(gpu, make):
    _per_step(make(), 26)


@pytest.mark.parametrize("make", [W.c5_box257, W.ocean385, W.c2_box513], ids=["box257", "box385", "box513"])
def test_step_parity_large(gpu, make):
    _per_step(make(), 3)


def test_trajectory_box129_10days(gpu):
    """C1: 240 leapfrog steps of 1 h, both sides evolving independently."""
    w = W.c1_box129()
    qg, o, (ent, wekp, txis, txin) = _setup(w)
    worst = 0.0
    for nt in range(1, w.nsteps + 1):
        qg.step(w.tdt)
        o.step(w.tdt, ent, wekp, txis, txin)
        if (nt - 1) % w.supp_every == 0:
            qg.suppress()
            o.suppress()
        if nt % 24 == 0 or nt == w.nsteps:
            gq, _, gp, _ = qg.get_state()
            oq, _, op, _ = o.get_state()
            worst = max(worst, rel_err(gq, oq), rel_err(gp, op))
    print(f"[traj] box129 10 days worst rel err {worst:.3e}", flush=True)
    assert worst <= TRAJ_TOL


def test_solve_inhomog_matches_oracle_and_residual(gpu):
    """solve_inhomog_eqn vs the oracle and the residual check of test_inhomog.F90:123-154."""
    rng = np.random.default_rng(11)
    for make in (W.c1_box129, W.atmos_channel):
        w = make()
        qg = w.make_qg()
        o = cr.Basin(w.params())
        rhs = rng.random((w.nyp, w.nxp))
        if w.cyclic:
            rhs[:, -1] = rhs[:, 0]
        rdm2 = qg.modes()[3]
        for m in range(1, w.nl + 1):
            g = qg.solve_inhomog(m, rhs)
            c = o.solve_inhomog(m, rhs)
            assert rel_err(g, c) <= 1e-12, (w.name, m)
            a = 1.0 / w.dx ** 2
            dx2 = np.zeros_like(g)
            dy2 = np.zeros_like(g)
            dx2[:, 1:-1] = a * (g[:, :-2] + g[:, 2:] - 2 * g[:, 1:-1])
            if w.cyclic:   # column nxp copies column 1 (numerics.F90:442-443)
                dx2[:, 0] = a * (g[:, -2] + g[:, 1] - 2 * g[:, 0])
                dx2[:, -1] = dx2[:, 0]
            dy2[1:-1, :] = a * (g[:-2, :] + g[2:, :] - 2 * g[1:-1, :])
            res = dx2 + dy2 - rdm2[m - 1] * g
            sl = (slice(1, -1), slice(None)) if w.cyclic else (slice(1, -1), slice(1, -1))
            r = np.abs(res[sl] - rhs[sl]).max() / np.abs(rhs[sl]).max()
            assert r <= 1e-11, (w.name, m, r)


def test_modes_match_oracle(gpu):
    for make in (W.c1_box129, W.atmos_channel, lambda: W.small_box(nl=4)):
        w = make()
        qg = w.make_qg()
        o = cr.Basin(w.params())
        ga, gl2m, gm2l, grd = qg.modes()
        oa, ol2m, om2l, ord_ = o.modes()
        assert rel_err(ga, oa) == 0.0
        assert rel_err(grd, ord_) <= 1e-13
        # eigenvector scaling and sign are free; each mode's projector ctl2m(:,m) ctm2l(m,:) is not
        for m in range(w.nl):
            assert rel_err(np.outer(gl2m[:, m], gm2l[m, :]), np.outer(ol2m[:, m], om2l[m, :])) <= 1e-12


def test_mass_constraint_identity_513(gpu):
    """Box: dz_sign (int p_k - int p_k+1) dA = dpi(k) after every step (homog.F90:345-364)."""
    w = W.c2_box513()
    qg = w.make_qg()
    o = cr.Basin(w.params())    # only its trapezoid integral is used
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    ent, wekp, txis, txin = w.forcing()
    qg.set_forcing(ent, wekp, txis, txin)
    for nt in range(1, 31):
        qg.step(w.tdt)
        if (nt - 1) % w.supp_every == 0:
            qg.suppress()
    _, _, gp, _ = qg.get_state()
    dpi = qg.get_constr()[: w.nl - 1]
    ints = [o.int_P_dA(gp[k]) for k in range(w.nl)]
    for k in range(w.nl - 1):
        lhs = w.dz_sign * (ints[k] - ints[k + 1])
        assert abs(lhs - dpi[k]) <= 1e-10 * max(abs(dpi[k]), abs(ints[k]) + abs(ints[k + 1])), k


def test_graph_run_equals_eager_and_is_deterministic(gpu):
    w = W.c2_box513()
    p = w.initial_p()
    ent, wekp, txis, txin = w.forcing()
    outs = []
    for mode in ("eager", "graph", "graph"):
        qg = w.make_qg()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        qg.set_forcing(ent, wekp, txis, txin)
        if mode == "eager":
            for nt in range(1, 101):
                qg.step(w.tdt)
                if (nt - 1) % w.supp_every == 0:
                    qg.suppress()
        else:
            qg.run(w.tdt, 100, 0, w.supp_every)
            qg.sync()
        outs.append(qg.get_state())
        qg.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    for a, b in zip(outs[1], outs[2]):
        assert np.array_equal(a, b)


def test_graph_replay_sees_new_channel_forcing(gpu):
    """A graph captured under one forcing must use the forcing set later
    (txis/txin and the layer-2 entrainment switch are read at run time)."""
    w = W.small_channel(33, 17, 3, seed=4)
    p = w.initial_p()
    _, wekp, txis, txin = w.forcing()
    rng = np.random.default_rng(9)
    ent_a = 1e-7 * rng.standard_normal((1, w.nyp, w.nxp))
    ent_a[..., -1] = ent_a[..., 0]
    ent_b = np.concatenate([ent_a, 1e-7 * rng.standard_normal((1, w.nyp, w.nxp))])
    ent_b[..., -1] = ent_b[..., 0]
    outs = []
    for mode in ("eager", "graph", "unchanged"):
        qg = w.make_qg()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        qg.set_forcing(ent_a, wekp, txis, txin)
        if mode == "eager":
            for nt in range(1, 5):
                qg.step(w.tdt)
        else:
            qg.run(w.tdt, 4, 0, 0)
        if mode != "unchanged":
            qg.set_forcing(ent_b, wekp, txis + 0.3, txin - 0.2)
        if mode == "eager":
            for nt in range(5, 9):
                qg.step(w.tdt)
        else:
            qg.run(w.tdt, 4, 4, 0)
        qg.sync()
        outs.append(qg.get_state())
        qg.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    assert not np.array_equal(outs[1][2], outs[2][2]), "the forcing change had no effect"


def test_kernel_times_advances_one_step(gpu):
    w = W.small_box(65, 3, seed=5)
    p = w.initial_p()
    ent, wekp, txis, txin = w.forcing()
    outs = []
    for timed in (False, True):
        qg = w.make_qg()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        qg.set_forcing(ent, wekp, txis, txin)
        qg.step(w.tdt)
        if timed:
            kt = qg.kernel_times(w.tdt, 4)
            assert set(kt) == {"k_vort", "k_fwd", "k_ysolve", "k_inv"}
            assert all(kt[k] > 0 for k in ("k_vort", "k_fwd", "k_ysolve", "k_inv"))
        else:
            qg.step(w.tdt)
        qg.step(w.tdt)
        outs.append((qg.get_state(), qg.get_constr()))
        qg.close()
    for a, b in zip(outs[0][0], outs[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(outs[0][1], outs[1][1])


def test_set_get_state_roundtrip_and_torch(gpu):
    import torch

    w = W.small_box()
    qg = w.make_qg()
    rng = np.random.default_rng(3)
    fields = [rng.standard_normal(qg.shape) for _ in range(4)]
    qg.set_state(*fields)
    for a, b in zip(qg.get_state(), fields):
        assert np.array_equal(a, b)
    t = [torch.from_numpy(f).cuda() for f in fields[::-1]]
    qg.set_state(*t)
    back = qg.get_state_torch()
    for a, b in zip(back, t):
        assert torch.equal(a, b)
    assert np.array_equal(qg.get_layer(2, 2), fields[::-1][2][1])


def test_cyclic_rejects_inconsistent_columns(gpu):
    from openqg_amd import OqgError

    w = W.small_channel()
    qg = w.make_qg()
    bad = np.zeros(qg.shape)
    bad[0, 3, 0] = 1.0
    with pytest.raises(OqgError):
        qg.set_state(p=bad)


def test_planned_run_equals_eager(gpu):
    """oqg_plan_run captures the head and tail stretches of a run; the planned run
    replays them and must equal stepping one step at a time, bitwise.  Planning
    itself runs nothing."""
    w = W.small_box(65, 3, seed=6)
    p = w.initial_p()
    ent, wekp, txis, txin = w.forcing()
    for n, s0, supp in ((20, 5, 25), (7, 0, 4), (33, 3, 4), (1, 0, 25), (61, 10, 25), (9, 2, 0)):
        outs = []
        for mode in ("eager", "planned"):
            qg = w.make_qg()
            qg.set_state(p=p, pm=p)
            qg.init_state()
            qg.set_forcing(ent, wekp, txis, txin)
            for nt in range(1, s0 + 1):      # the steps before the run
                qg.step(w.tdt)
                if supp and (nt - 1) % supp == 0:
                    qg.suppress()
            if mode == "eager":
                for nt in range(s0 + 1, s0 + n + 1):
                    qg.step(w.tdt)
                    if supp and (nt - 1) % supp == 0:
                        qg.suppress()
            else:
                before = qg.get_state()
                qg.plan_run(w.tdt, n, s0, supp)
                qg.sync()
                for a, b in zip(qg.get_state(), before):
                    assert np.array_equal(a, b), "planning changed the state"
                qg.run(w.tdt, n, s0, supp)
            qg.sync()
            outs.append((qg.get_state(), qg.get_constr()))
            qg.close()
        for a, b in zip(outs[0][0], outs[1][0]):
            assert np.array_equal(a, b), (n, s0, supp)
        assert np.array_equal(outs[0][1], outs[1][1])


def test_unaligned_stretches_equal_eager(gpu):
    """oqg_run over stretches that are not whole suppression periods (the
    advances of a checkpointed sweep), chained from every phase and both
    parities, must equal stepping one step at a time, bitwise, and so must
    a second context running the same chain."""
    w = W.small_box(65, 3, seed=11)
    p = w.initial_p()
    ent, wekp, txis, txin = w.forcing()
    # (step0, n) pairs chained one after another: every phase of supp = 25,
    # odd and even lengths, single steps, stretches across a whole period
    runs = [(0, 3), (3, 1), (4, 30), (34, 47), (81, 2), (83, 19), (102, 64), (166, 5)]
    supp = w.supp_every
    ref = None
    for mode in ("eager", "run", "run again"):
        qg = w.make_qg()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        qg.set_forcing(ent, wekp, txis, txin)
        for s0, n in runs:
            if mode == "eager":
                for nt in range(s0 + 1, s0 + n + 1):
                    qg.step(w.tdt)
                    if (nt - 1) % supp == 0:
                        qg.suppress()
            else:
                qg.run(w.tdt, n, s0, supp)
        qg.sync()
        out = (qg.get_state(), qg.get_constr())
        qg.close()
        if ref is None:
            ref = out
            continue
        for a, b in zip(ref[0], out[0]):
            assert np.array_equal(a, b), mode
        assert np.array_equal(ref[1], out[1]), mode


@pytest.mark.parametrize("make", [lambda: W.small_box(65, 3), lambda: W.small_channel(97, 49, 3)],
                         ids=["box65", "chan97x49"])
def test_tile_heights_agree_and_tune_restores_state(gpu, make):
    """The vorticity kernel's tile height (4, 8, 16 rows) leaves every point's
    arithmetic as it is: without entrainment the runs agree bitwise; with it
    only the summation order of its layer-1 integral differs.  oqg_tune keeps
    the state it times from and sets the fastest height."""
    w = make()
    print(f"[parity] tile heights, {w.name}", flush=True)
    for ent_amp in (0.0, 1e-6):
        runs = {}
        for rows in (4, 8, 16):
            qg, _, _ = _setup(w, ent_amp=ent_amp)
            qg.tile_rows = rows
            assert qg.tile_rows == rows
            qg.run(w.tdt, 7, 0, 3)
            runs[rows] = qg.get_state()
            qg.close()
        for rows in (8, 16):
            for a, b in zip(runs[rows], runs[4]):
                if ent_amp == 0.0:
                    assert np.array_equal(a, b), rows
                else:
                    assert rel_err(a, b) <= 1e-13, rows
    qg, _, _ = _setup(w)
    qg.run(w.tdt, 3, 0, 3)
    before = qg.get_state()
    rows, ms = qg.tune(w.tdt, trial_steps=6, supp_every=3)
    assert rows in (4, 8, 16) and qg.tile_rows == rows and all(t > 0 for t in ms.values())
    for a, b in zip(qg.get_state(), before):
        assert np.array_equal(a, b)
    qg.close()
