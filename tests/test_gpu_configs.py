"""The measured configurations at their full sizes and lengths, on the GPU.

- 513^2 x 3 box, 30 days (2160 steps of dt = 1200 s): every step checked
  against the CPU oracle from the GPU's own state (per-step bar 1e-12 on q and
  p, 1e-9 on the constraint scalars).
- The 64-member uneven ensemble (65^2..513^2, 2-4 layers): the batched run
  equals each member stepped alone, bitwise, for every member (so for every
  grid and layer class), in both batched forms; a 385^2 and a 513^2 member are
  also checked step by step against the oracle while batched.
- The coupled pair, 385^2 ocean + 97x49 atmosphere channel, through the run
  loop for 30 days (dta = 200 s, nstr = 6): equal, bitwise, to the two basins
  stepped one step at a time with the reference's cadence
  (src/openqg.F90:483-509), with per-step oracle checks of both basins along
  the way.
"""
import numpy as np
import pytest
import torch

from oracle import cpuref as cr
from openqg_amd import workloads as W
from openqg_amd.driver import Basin, Clock, Valids, run
from openqg_amd.ensemble import Ensemble, EnsembleGroups

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12


def _maxrel(a, b):
    """max |a - b| / max |b| with torch's threaded CPU kernels (the fields are large)."""
    ta, tb = torch.from_numpy(a), torch.from_numpy(b)
    den = float(tb.abs().max())
    return float((ta - tb).abs().max()) / (den if den > 0 else 1.0)


def _constr_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b), np.abs(b).max() * 1e-3 + 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def _start(w):
    qg = w.make_qg()
    w.start(qg)
    fo = w.forcing()
    qg.set_forcing(*fo)
    return qg, fo


def _check_step(qg, o, w, fo, nt, supp):
    """One step of qg and of the oracle from qg's current state; returns the errors."""
    ent, wekp, txis, txin = fo
    S = qg.get_state()
    o.set_state(*S)
    o.set_constr(qg.get_constr())
    qg.step(w.tdt)
    o.step(w.tdt, ent, wekp, txis, txin)
    if supp and (nt - 1) % supp == 0:
        qg.suppress()
        o.suppress()
    gq, _, gp, _ = qg.get_state()
    oq, _, op, _ = o.get_state()
    return _maxrel(gq, oq), _maxrel(gp, op), _constr_err(qg.get_constr(), o.get_constr())


def test_box513_30_days_every_step_against_the_oracle(gpu):
    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    w = W.c2_box513()
    qg, fo = _start(w)
    o = cr.Basin(w.params())
    print(f"[c2] {w.name}: {w.nsteps} steps, per-step check", flush=True)
    worst = [0.0, 0.0, 0.0]
    for nt in range(1, w.nsteps + 1):
        eq, ep, ec = _check_step(qg, o, w, fo, nt, w.supp_every)
        worst = [max(a, b) for a, b in zip(worst, (eq, ep, ec))]
        assert eq <= STEP_TOL and ep <= STEP_TOL, f"step {nt}: q {eq:.3e} p {ep:.3e}"
        assert ec <= 1e-9, f"step {nt}: constraints {ec:.3e}"
        if nt % 360 == 0:
            print(f"[c2]   step {nt}: worst q {worst[0]:.2e} p {worst[1]:.2e} constraints {worst[2]:.2e}", flush=True)
    assert np.isfinite(qg.get_state()[2]).all()
    qg.close()


# ------------------------------------------------------------------ C3


def _ensemble_members(ws):
    return [_start(w)[0] for w in ws]


def _alone(ws, nsteps, supp):
    out = []
    for w, n in zip(ws, nsteps):
        qg, _ = _start(w)
        for nt in range(1, n + 1):
            qg.step(w.tdt)
            if supp and (nt - 1) % supp == 0:
                qg.suppress()
        out.append((qg.get_state(), qg.get_constr()))
        qg.close()
    return out


@pytest.fixture(scope="module")
def uneven():
    ws = W.uneven_ensemble(64)
    nsteps = [5 + (i % 7) for i in range(64)]     # uneven lengths: members leave the batch at different steps
    supp = 4
    ref = _alone(ws, nsteps, supp)
    return ws, nsteps, supp, ref


def test_uneven_ensemble_covers_every_class(gpu, uneven):
    ws = uneven[0]
    classes = {(w.nxp, w.nl) for w in ws}
    assert classes == {(n, nl) for n in W.ENS_SIZES for nl in (2, 3, 4)}, sorted(classes)


@pytest.mark.parametrize("form", ["one_batch", "groups_and_solo"])
def test_uneven_ensemble_64_batched_equals_alone(gpu, uneven, form):
    ws, nsteps, supp, ref = uneven
    qgs = _ensemble_members(ws)
    if form == "one_batch":
        ens = Ensemble(qgs)
    else:
        ens = EnsembleGroups(qgs, 2, solo_points=513 * 513)
    print(f"[c3] 64 members, {form}", flush=True)
    ens.run([w.tdt for w in ws], nsteps, 0, supp)
    ens.sync()
    for i, (qg, (st, con)) in enumerate(zip(qgs, ref)):
        for name, a, b in zip(("q", "qm", "p", "pm"), qg.get_state(), st):
            assert np.array_equal(a, b), f"member {i} ({ws[i].name}) {name}"
        assert np.allclose(qg.get_constr(), con, rtol=1e-10, atol=0.0)
    ens.close()
    for qg in qgs:
        qg.close()


def test_batched_large_members_against_the_oracle(gpu):
    """Members of 385^2 and 513^2 inside a batch with small ones, one batched step at a time,
    each step checked against the oracle from the member's own state."""
    ws = W.uneven_ensemble(64)
    pick = [next(i for i, w in enumerate(ws) if w.nxp == n) for n in (385, 513)]
    small = [i for i, w in enumerate(ws) if w.nxp <= 129][:4]
    sel = [ws[i] for i in pick + small]
    qgs, fos = zip(*[_start(w) for w in sel])
    oracles = [cr.Basin(w.params()) for w in sel[:2]]
    ens = Ensemble(list(qgs))
    supp = 4
    print(f"[c3] oracle checks of {[w.name for w in sel[:2]]} in a batch of {len(sel)}", flush=True)
    for nt in range(1, 11):
        for qg, o in zip(qgs[:2], oracles):
            o.set_state(*qg.get_state())
            o.set_constr(qg.get_constr())
        ens.run([w.tdt for w in sel], 1, nt - 1, supp)
        ens.sync()
        for w, qg, o, fo in zip(sel[:2], qgs[:2], oracles, fos[:2]):
            o.step(w.tdt, *fo)
            if (nt - 1) % supp == 0:
                o.suppress()
            gq, _, gp, _ = qg.get_state()
            oq, _, op, _ = o.get_state()
            eq, ep = _maxrel(gq, oq), _maxrel(gp, op)
            assert eq <= STEP_TOL and ep <= STEP_TOL, (w.name, nt, eq, ep)
            assert _constr_err(qg.get_constr(), o.get_constr()) <= 1e-9
    ens.close()


# ------------------------------------------------------------------ C4


def test_coupled_pair_30_days_through_the_run_loop(gpu, tmp_path):
    wo, wa = W.ocean385(), W.atmos_channel()
    clk = Clock(dta=200.0, nstr=6, tini=0.0, trun=30.0 / 365.0, resday=10.0)
    assert clk.dto == wo.dt and clk.dta == wa.dt and clk.ntsrun == 12960
    print(f"[c4] {wo.name} + {wa.name}: {clk.ntsrun} atmosphere steps", flush=True)
    o1, _ = _start(wo)
    a1, _ = _start(wa)
    nt = run(clk, str(tmp_path), ocn=Basin("ocean", o1, Valids(1440, 1e12, 1e3)),
             atm=Basin("atmos", a1, Valids(1440, 1e12, 1e3)))
    assert nt == clk.nsteps

    o2, fo_o = _start(wo)
    a2, fo_a = _start(wa)
    oo, oa = cr.Basin(wo.params()), cr.Basin(wa.params())
    checks = {1, 2, 6, 7, 3001, 6000, 12955}
    worst = 0.0
    for t in range(1, clk.nsteps + 1):
        if t in checks:
            eq, ep, ec = _check_step(a2, oa, wa, fo_a, t, 100)
        else:
            a2.step(clk.tdta)
            if (t - 1) % 100 == 0:
                a2.suppress()
        if (t - 1) % clk.nstr == 0:
            to = (t - 1) // clk.nstr + 1
            if t in checks:
                eq2, ep2, ec2 = _check_step(o2, oo, wo, fo_o, to, 25)
                assert max(eq2, ep2) <= STEP_TOL and ec2 <= 1e-9, ("ocean", t, eq2, ep2, ec2)
                worst = max(worst, eq2, ep2)
            else:
                o2.step(clk.tdto)
                if (to - 1) % 25 == 0:
                    o2.suppress()
        if t in checks:
            assert max(eq, ep) <= STEP_TOL and ec <= 1e-9, ("atmos", t, eq, ep, ec)
            worst = max(worst, eq, ep)
    print(f"[c4] oracle spot checks worst {worst:.2e}", flush=True)
    for name, x, y in (("ocean", o1, o2), ("atmos", a1, a2)):
        for f, u, v in zip(("q", "qm", "p", "pm"), x.get_state(), y.get_state()):
            assert u.tobytes() == v.tobytes(), f"{name} {f}"
            assert np.isfinite(u).all()
        np.testing.assert_array_equal(x.get_constr(), y.get_constr())


# ------------------------------------------------------------------ C5


def _c5_setup():
    from openqg_amd.forcing import double_gyre_tau

    w = W.c5_box257()
    qg = w.make_qg()
    qg.set_state(p=w.initial_p(), pm=w.initial_p())
    qg.init_state()
    q0 = qg.get_state()[0]
    tx, ty = double_gyre_tau(w.nyp, w.nxp, w.dx, w.tau0)
    ty = 0.3 * np.roll(tx, 3, axis=0).T.copy()    # a meridional component too
    return w, qg, q0, tx, ty


def _interior(a):
    b = np.zeros_like(a)
    b[:, 1:-1, 1:-1] = a[:, 1:-1, 1:-1]
    return b


def _unit_directions(q0, tx, ty, n=16, seed=5):
    """n seeded directions of the joint control (taux, tauy, interior q0), each block scaled to the
    norm of that block of the control, so that eps is relative to ||x||."""
    ntau = float(np.sqrt(np.sum(tx * tx) + np.sum(ty * ty)))
    nq = float(np.sqrt(np.sum(_interior(q0) ** 2)))
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dx_, dy_ = rng.standard_normal(tx.shape), rng.standard_normal(ty.shape)
        dq = _interior(rng.standard_normal(q0.shape))
        nt = np.sqrt(np.sum(dx_ * dx_) + np.sum(dy_ * dy_))
        out.append((dq * (nq / np.sqrt(np.sum(dq * dq))), dx_ * (ntau / nt), dy_ * (ntau / nt)))
    return out


def _fd(qg, w, q0, tx, ty, d, nsteps, eps):
    from openqg_amd.adjoint import objective

    dq, dx_, dy_ = d
    jp = objective(qg, q0 + eps * dq, tx + eps * dx_, ty + eps * dy_, w.tdt, nsteps, w.supp_every)
    jm = objective(qg, q0 - eps * dq, tx - eps * dx_, ty - eps * dy_, w.tdt, nsteps, w.supp_every)
    return (jp - jm) / (2 * eps)


def test_gradient_257_ten_days_sixteen_directions(gpu):
    """257^2 x 3, 720 steps (10 days), revolve with 32 snapshots: central differences with
    eps = 1e-6 ||x|| in 16 seeded unit directions agree with the adjoint to 1e-6."""
    from openqg_amd.adjoint import gradient

    w, qg, q0, tx, ty = _c5_setup()
    n = 720
    J, gx, gy, gq, fs = gradient(qg, q0, tx, ty, w.tdt, n, w.supp_every, snapshots=32)
    worst = 0.0
    for k, d in enumerate(_unit_directions(q0, tx, ty)):
        fd = _fd(qg, w, q0, tx, ty, d, n, 1e-6)
        ad = float(np.sum(gq * d[0]) + np.sum(gx * d[1]) + np.sum(gy * d[2]))
        err = abs(fd - ad) / abs(fd)
        worst = max(worst, err)
        assert err <= 1e-6, (k, fd, ad, err)
    print(f"[c5] {n} steps, revolve-32 ({fs} forward steps): worst rel err {worst:.2e}", flush=True)


def test_gradient_257_thirty_days_sixteen_directions(gpu):
    """The full gradient workload: 257^2 x 3, 2160 steps (30 days), revolve with 32 snapshots,
    16 seeded unit directions.  Over 30 days the flow's sensitivity is far from linear at
    eps = 1e-6 ||x|| (the central difference's eps^2 term is visible), so the check is that the
    differences converge to the adjoint: plain differences at eps = 1e-6 agree to a few percent,
    and Richardson's extrapolation of eps = 1e-6 and 5e-7, which cancels the eps^2 term, agrees
    much more closely."""
    from openqg_amd.adjoint import gradient

    w, qg, q0, tx, ty = _c5_setup()
    n = w.nsteps
    J, gx, gy, gq, fs = gradient(qg, q0, tx, ty, w.tdt, n, w.supp_every, snapshots=32)
    assert n == 2160 and fs == 5854
    e1, e2 = [], []
    for k, d in enumerate(_unit_directions(q0, tx, ty)):
        f1 = _fd(qg, w, q0, tx, ty, d, n, 1e-6)
        f2 = _fd(qg, w, q0, tx, ty, d, n, 5e-7)
        rich = (4.0 * f2 - f1) / 3.0
        ad = float(np.sum(gq * d[0]) + np.sum(gx * d[1]) + np.sum(gy * d[2]))
        e1.append(abs(f1 - ad) / abs(ad))
        e2.append(abs(rich - ad) / abs(ad))
        print(f"[c5]   dir {k:2d}: adj {ad:+.10e} fd(1e-6) {e1[-1]:.2e} richardson {e2[-1]:.2e}", flush=True)
    print(f"[c5] 2160 steps: worst fd {max(e1):.2e}, worst richardson {max(e2):.2e}", flush=True)
    assert max(e1) <= 3e-2
    assert max(e2) <= 1e-3 and np.median(e2) <= 2e-5
    assert np.median(e2) <= 0.1 * np.median(e1)


def test_gradient_batch_257_sixteen_members_30_days(gpu):
    """The gradient workload's batch: 16 members of 257^2 x 3 (their own PV seed and wind
    amplitude), 2160 steps, revolve-32, concurrent on their own streams: each member's gradient
    equals the one computed alone, bitwise."""
    import dataclasses

    from openqg_amd.adjoint import gradient, gradient_batch
    from openqg_amd.forcing import double_gyre_tau

    base = W.c5_box257()
    members = [dataclasses.replace(base, seed=base.seed + i, tau0=base.tau0 * (1.0 + 0.05 * i)) for i in range(16)]
    qgs, q0s, txs, tys = [], [], [], []
    for w in members:
        qg = w.make_qg()
        qg.set_state(p=w.initial_p(), pm=w.initial_p())
        qg.init_state()
        q0s.append(qg.get_state()[0])
        tx, ty = double_gyre_tau(w.nyp, w.nxp, w.dx, w.tau0)
        txs.append(tx)
        tys.append(ty)
        qgs.append(qg)
    n = base.nsteps
    got = gradient_batch(qgs, q0s, txs, tys, base.tdt, n, base.supp_every, 32)
    for i in (0, 7, 15):
        ref = gradient(qgs[i], q0s[i], txs[i], tys[i], base.tdt, n, base.supp_every, 32)
        assert ref[0] == got[i][0]
        for k in (1, 2, 3):
            assert np.array_equal(ref[k], got[i][k]), (i, k)
    assert len({g[0] for g in got}) == 16
    for qg in qgs:
        qg.close()
