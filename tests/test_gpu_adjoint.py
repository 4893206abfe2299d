"""The hand-written adjoint of step_qg (box) against finite differences.

One step of the model is a quadratic map of the state (q, qm, p, pm) and the
Ekman pumping (bilinear Jacobian, everything else linear), so a central
difference of it is exact up to rounding: the dot-product test
<lam, (F(S + e d) - F(S - e d)) / 2e> == <F'(S)^T lam, d> holds to ~1e-12
relative, separately for each input component so that no identity part of
the leapfrog hides a wrong transpose.  Over many steps the map is no longer
quadratic and the gradient is checked against central differences of the
objective, on the GPU model and on the CPU oracle (the reference restatement).
"""
import numpy as np
import pytest

from oracle import cpuref as cr
from openqg_amd import workloads as W
from openqg_amd.adjoint import AdjointState, ekei1, gradient, make_autograd, objective
from openqg_amd.forcing import double_gyre_tau, ekman_wekp

pytestmark = pytest.mark.gpu

DOT_TOL = 1e-9


def _rms(a):
    return float(np.sqrt(np.mean(a * a))) or 1.0


def _prepared(w, nsteps):
    qg = w.make_qg()
    qg.set_state(p=w.initial_p(), pm=w.initial_p())
    qg.init_state()
    ent, wekp, txis, txin = w.forcing()
    qg.set_forcing(ent, wekp, txis, txin)
    for nt in range(1, nsteps + 1):
        qg.step(w.tdt)
        if (nt - 1) % w.supp_every == 0:
            qg.suppress()
    return qg, ent, wekp


def _fstep(qg, slot_adj, S, wekp, ent, tdt, suppressed):
    slot_adj.load(0)
    qg.set_state(*S)
    qg.set_forcing(ent, wekp, 0.0, 0.0)
    qg.step(tdt)
    if suppressed:
        qg.suppress()
    return qg.get_state()


@pytest.mark.parametrize("suppressed", [False, True], ids=["plain", "suppressed"])
def test_one_step_dot_product(gpu, suppressed):
    w = W.small_box(33, 3)
    qg, ent, wekp = _prepared(w, 3)
    adj = AdjointState(qg)
    adj.alloc(1)
    adj.save(0)
    S = qg.get_state()
    scale = [_rms(a) for a in S]
    # the step is affine in wekp, so any size of its perturbation keeps the
    # central difference exact; 1e3 x rms(wekp) gives a response of the size
    # of the state parts', well above the rounding of the differences
    wscale = 1e3 * _rms(wekp)
    rng = np.random.default_rng(11 + suppressed)
    lam_all = [rng.standard_normal(qg.shape) / s for s in scale]
    lam_dyn = [lam_all[0], np.zeros(qg.shape), lam_all[2], np.zeros(qg.shape)]
    print(f"[adjoint] one-step dot product, {w.name}, suppressed={suppressed}", flush=True)
    for lam_name, lam in (("all", lam_all), ("q,p", lam_dyn)):
        # adjoint: lam is the adjoint of S_{n+1}; with S_{n+1} loaded, adj.set
        # addresses it, and the step back is taken with S_n loaded again
        _fstep(qg, adj, S, wekp, ent, w.tdt, suppressed)
        adj.set(*lam)
        adj.zero_wekp()
        adj.load(0)
        adj.step(w.tdt, suppressed)
        lS = adj.get()
        lw = adj.wekp()
        for comp in ("q", "qm", "p", "pm", "wekp", "all"):
            d = [np.zeros(qg.shape) for _ in range(4)]
            dw = np.zeros_like(wekp)
            for c in range(4):
                if comp in (("q", "qm", "p", "pm")[c], "all"):
                    d[c] = scale[c] * rng.standard_normal(qg.shape)
            if comp in ("wekp", "all"):
                dw = wscale * rng.standard_normal(wekp.shape)
            eps = 1e-3
            Fp = _fstep(qg, adj, [s + eps * x for s, x in zip(S, d)], wekp + eps * dw, ent, w.tdt, suppressed)
            Fm = _fstep(qg, adj, [s - eps * x for s, x in zip(S, d)], wekp - eps * dw, ent, w.tdt, suppressed)
            fd = sum(float(np.sum(l * (a - b))) for l, a, b in zip(lam, Fp, Fm)) / (2 * eps)
            ad = sum(float(np.sum(l * x)) for l, x in zip(lS, d)) + float(np.sum(lw * dw))
            err = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
            print(f"[adjoint]   lam={lam_name:4s} d={comp:5s} fd={fd:+.12e} adj={ad:+.12e} rel={err:.2e}", flush=True)
            assert err <= DOT_TOL, (lam_name, comp, fd, ad)
    qg.set_forcing(ent, wekp, 0.0, 0.0)


def _controls(w, qg):
    """Initial interior PV and the double-gyre kinematic stress of workload w."""
    qg.set_state(p=w.initial_p(), pm=w.initial_p())
    qg.init_state()
    q0 = qg.get_state()[0]
    taux, tauy = double_gyre_tau(w.nyp, w.nxp, w.dx, w.tau0)
    # a meridional component too, so both stress adjoints are exercised
    tauy = 0.3 * np.roll(taux, 3, axis=0).T.copy() if w.nxp == w.nyp else tauy
    return q0, taux, tauy


def _fd(qg, q0, tx, ty, d, w, nsteps, supp, rel=1e-5):
    dq, dx, dy = d
    eps = rel
    jp = objective(qg, q0 + eps * dq, tx + eps * dx, ty + eps * dy, w.tdt, nsteps, supp)
    jm = objective(qg, q0 - eps * dq, tx - eps * dx, ty - eps * dy, w.tdt, nsteps, supp)
    return (jp - jm) / (2 * eps)


def _interior(a):
    b = np.zeros_like(a)
    b[:, 1:-1, 1:-1] = a[:, 1:-1, 1:-1]
    return b


@pytest.mark.parametrize("n", [33, 65])
def test_device_ke_seed_matches_host_restatement(gpu, n):
    """oqg_adj_seed_ke: J = ekei(1) from the device's row sums and dJ/dp_1 by
    the gather form of the derivative transposes, against the host
    restatement ekei1 (whose gradient is pinned by central differences on the
    CPU); the other adjoint fields are zeroed."""
    w = W.small_box(n, 3, seed=3)
    qg, _, _ = _prepared(w, 5)
    adj = AdjointState(qg)
    adj.set(lq=np.ones(qg.shape), lqm=np.ones(qg.shape), lp=np.ones(qg.shape), lpm=np.ones(qg.shape))
    J = adj.seed_ke(1000.0)
    lq, lqm, lp, lpm = adj.get()
    p = qg.get_state()[2]
    Jh, gh = ekei1(p[0], w.dx, w.dx, w.fnot, qg.bcco, 1000.0, w.h[0])
    assert abs(J - Jh) <= 1e-13 * abs(Jh)
    assert np.abs(lp[0] - gh).max() <= 1e-13 * np.abs(gh).max()
    assert not lp[1:].any() and not lq.any() and not lqm.any() and not lpm.any()


def test_gradient_matches_finite_differences(gpu):
    w = W.small_box(33, 3)
    nsteps, supp = 10, 4
    qg = w.make_qg()
    q0, tx, ty = _controls(w, qg)
    J, gx, gy, gq, fsteps = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp)
    J2 = objective(qg, q0, tx, ty, w.tdt, nsteps, supp)
    assert fsteps == nsteps
    assert J > 0 and abs(J - J2) <= 1e-13 * J, (J, J2)
    assert np.all(gq[:, 0, :] == 0) and np.all(gq[:, -1, :] == 0)
    assert np.all(gq[:, :, 0] == 0) and np.all(gq[:, :, -1] == 0)
    rng = np.random.default_rng(5)
    dirs = {
        "taux": (np.zeros_like(q0), _rms(tx) * rng.standard_normal(tx.shape), np.zeros_like(ty)),
        "tauy": (np.zeros_like(q0), np.zeros_like(tx), _rms(tx) * rng.standard_normal(ty.shape)),
        "q0": (_interior(_rms(q0) * rng.standard_normal(q0.shape)), np.zeros_like(tx), np.zeros_like(ty)),
    }
    print(f"[adjoint] gradient vs central differences, {w.name}, {nsteps} steps, J={J:.6e}", flush=True)
    for name, (dq, dx_, dy_) in dirs.items():
        fd = _fd(qg, q0, tx, ty, (dq, dx_, dy_), w, nsteps, supp)
        ad = float(np.sum(gq * dq) + np.sum(gx * dx_) + np.sum(gy * dy_))
        err = abs(fd - ad) / abs(fd)
        print(f"[adjoint]   {name:5s} fd={fd:+.10e} adj={ad:+.10e} rel={err:.2e}", flush=True)
        assert err <= 1e-6, (name, fd, ad)


def test_gradient_sixteen_unit_directions(gpu):
    """Central differences in 16 seeded unit directions (seed 5) of the joint
    control (taux, tauy, interior q0), each block scaled to its norm, with
    eps = 1e-6 ||x||: dJ/de at e = 0 of J(x + e ||x_b|| d_b)."""
    w = W.small_box(33, 3)
    nsteps, supp = 10, 4
    qg = w.make_qg()
    q0, tx, ty = _controls(w, qg)
    J, gx, gy, gq, _ = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp)
    ntau = float(np.sqrt(np.sum(tx * tx) + np.sum(ty * ty)))
    nq = float(np.sqrt(np.sum(_interior(q0) ** 2)))
    rng = np.random.default_rng(5)
    worst = 0.0
    for k in range(16):
        dx_, dy_ = rng.standard_normal(tx.shape), rng.standard_normal(ty.shape)
        dq = _interior(rng.standard_normal(q0.shape))
        nt = np.sqrt(np.sum(dx_ * dx_) + np.sum(dy_ * dy_))
        dx_, dy_ = dx_ * (ntau / nt), dy_ * (ntau / nt)
        dq = dq * (nq / np.sqrt(np.sum(dq * dq)))
        fd = _fd(qg, q0, tx, ty, (dq, dx_, dy_), w, nsteps, supp, rel=1e-6)
        ad = float(np.sum(gq * dq) + np.sum(gx * dx_) + np.sum(gy * dy_))
        err = abs(fd - ad) / abs(fd)
        worst = max(worst, err)
        assert err <= 1e-5, (k, fd, ad)
    print(f"[adjoint] 16 unit directions, {w.name}: worst rel err {worst:.2e}", flush=True)


def test_checkpointed_gradient_equals_stored(gpu):
    w = W.small_box(33, 3)
    nsteps, supp = 23, 4
    qg = w.make_qg()
    q0, tx, ty = _controls(w, qg)
    ref = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp, snapshots=None)
    for s in (1, 3, 5):
        got = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp, snapshots=s)
        print(f"[adjoint] revolve with {s} snapshots: {got[4]} forward steps for {nsteps}", flush=True)
        assert got[0] == ref[0]
        for a, b in zip(got[1:4], ref[1:4]):
            np.testing.assert_array_equal(a, b)
        assert got[4] > nsteps


def test_autograd_function(gpu):
    import torch

    w = W.small_box(33, 3)
    nsteps, supp = 6, 4
    qg = w.make_qg()
    q0, tx, ty = _controls(w, qg)
    J, gx, gy, gq, _ = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp)
    f = make_autograd(qg, w.tdt, nsteps, supp)
    t = [torch.tensor(a, requires_grad=True) for a in (tx, ty, q0)]
    out = f(*t)
    (2.0 * out).backward()
    assert out.detach().item() == J
    for a, ref in zip(t, (gx, gy, gq)):
        np.testing.assert_array_equal(a.grad.numpy(), 2.0 * ref)


def _oracle_objective(w, S0, c0, tx, ty, nsteps, supp, bcco, h1):
    o = cr.Basin(w.params())
    o.set_state(*S0)
    o.set_constr(c0)
    wekp, _, _ = ekman_wekp(w.nxp, w.nyp, w.dx, w.dx, w.fnot, False, tx, ty)
    ent = np.zeros((max(w.nl - 1, 1), w.nyp, w.nxp))
    for nt in range(1, nsteps + 1):
        o.step(w.tdt, ent, wekp)
        if (nt - 1) % supp == 0:
            o.suppress()
    return ekei1(o.get_state()[2][0], w.dx, w.dx, w.fnot, bcco, 1000.0, h1)[0]


def test_stress_gradient_matches_oracle_perturbation(gpu):
    """Reference gradient: central differences of the CPU oracle's model, from
    the same initial state, in random stress directions."""
    w = W.small_box(33, 3)
    nsteps, supp = 10, 4
    qg = w.make_qg()
    q0, tx, ty = _controls(w, qg)
    J, gx, gy, _, _ = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp)
    # the initial state of the run, after the inversion of q0
    qg.set_state(q=q0, qm=q0, p=np.zeros(qg.shape), pm=np.zeros(qg.shape))
    qg.set_constr(np.zeros(qg._lib.oqg_constr_len(qg._h)))
    AdjointState(qg).invert()
    S0, c0 = qg.get_state(), qg.get_constr()
    Jo = _oracle_objective(w, S0, c0, tx, ty, nsteps, supp, qg.bcco, w.h[0])
    assert abs(Jo - J) <= 1e-10 * J, (Jo, J)
    rng = np.random.default_rng(9)
    s = _rms(tx)
    for k in range(3):
        dx_, dy_ = s * rng.standard_normal(tx.shape), s * rng.standard_normal(ty.shape)
        eps = 1e-5
        jp = _oracle_objective(w, S0, c0, tx + eps * dx_, ty + eps * dy_, nsteps, supp, qg.bcco, w.h[0])
        jm = _oracle_objective(w, S0, c0, tx - eps * dx_, ty - eps * dy_, nsteps, supp, qg.bcco, w.h[0])
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(gx * dx_) + np.sum(gy * dy_))
        err = abs(fd - ad) / abs(fd)
        print(f"[adjoint] oracle direction {k}: fd={fd:+.10e} adj={ad:+.10e} rel={err:.2e}", flush=True)
        assert err <= 1e-6


def test_gradient_full_size_finite_differences(gpu):
    """257^2 three-layer box (the gradient workload's grid), 40 steps."""
    w = W.c5_box257()
    nsteps, supp = 40, 25
    qg = w.make_qg()
    q0, tx, ty = _controls(w, qg)
    J, gx, gy, gq, _ = gradient(qg, q0, tx, ty, w.tdt, nsteps, supp, snapshots=4)
    rng = np.random.default_rng(13)
    d = (_interior(_rms(q0) * rng.standard_normal(q0.shape)), _rms(tx) * rng.standard_normal(tx.shape),
         _rms(tx) * rng.standard_normal(ty.shape))
    fd = _fd(qg, q0, tx, ty, d, w, nsteps, supp)
    ad = float(np.sum(gq * d[0]) + np.sum(gx * d[1]) + np.sum(gy * d[2]))
    err = abs(fd - ad) / abs(fd)
    print(f"[adjoint] {w.name} {nsteps} steps: J={J:.6e} fd={fd:+.10e} adj={ad:+.10e} rel={err:.2e}", flush=True)
    assert err <= 1e-6
