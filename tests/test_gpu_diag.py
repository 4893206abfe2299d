"""Device diagnostics (oqg_diagnose) against the numpy restatement of
qg_diagno / scan_cfl / valids (oracle/monitor.py) on evolved states of a box
and a channel and on the 513^2 initial state."""
import numpy as np
import pytest

from oracle import monitor as M
from openqg_amd import workloads as W

pytestmark = pytest.mark.gpu

SUM_TOL = 1e-12   # sums in a different order than numpy's


def _check(qg, w, dt):
    q, _, p, _ = qg.get_state()
    d = qg.diagnose(dt)
    o = M.diagnose(q, p, w.dx, w.dx, w.fnot, qg.bcco, w.cyclic, w.h, w.gp, w.params()["dz_sign"], qg.rho, dt)
    for name in ("ekei", "pavg", "qavg", "etam", "et2m", "circ", "sfmin", "sfmax"):
        a, b = np.asarray(d[name]), np.asarray(o[name])
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= SUM_TOL * scale, (name, a, b)
    for name in ("pmin", "pmax", "qmin", "qmax", "umax", "vmax", "cflu", "cflv"):
        np.testing.assert_allclose(d[name], o[name], rtol=1e-15, atol=0, err_msg=name)
    assert abs(d["ctot"] - o["ctot"]) <= 1e-13 * max(abs(o["ctot"]), 1e-300)
    assert d["nonfinite"] == 0.0
    return d


@pytest.mark.parametrize("make,steps", [(W.small_box, 12), (W.small_channel, 12), (W.c2_box513, 0)],
                         ids=["box33", "chan33x17", "box513"])
def test_diagnostics_match_oracle(gpu, make, steps):
    w = make()
    qg = w.make_qg()
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    ent, wekp, txis, txin = w.forcing()
    qg.set_forcing(ent, wekp, txis, txin)
    print(f"[diag] {w.name} after {steps} steps", flush=True)
    if steps:
        qg.run(w.tdt, steps, 0, w.supp_every)
        qg.sync()
    d = _check(qg, w, w.dt)
    assert np.all(d["ekei"] > 0)


def test_nonfinite_values_are_counted(gpu):
    w = W.small_box(33, 3)
    qg = w.make_qg()
    qg.set_state(p=w.initial_p(), pm=w.initial_p())
    qg.init_state()
    q, qm, p, pm = qg.get_state()
    q[1, 5, 7] = np.nan
    p[2, 3, 3] = np.inf
    p[0, 0, 0] = -np.inf
    qg.set_state(q, qm, p, pm)
    assert qg.diagnose(w.dt)["nonfinite"] == 3.0


def test_continuity_check_matches_oracle(gpu):
    """check_continuity (qg_monitor.F90:571-604) on a channel evolved with
    entrainment: ermas and emfr from the device's layer integrals and the
    constraint's dpi against the oracle's int_P_dA of the same p."""
    w = W.small_channel()
    qg = w.make_qg()
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    ent, wekp, txis, txin = w.forcing()
    ent = 1e-6 * np.random.default_rng(w.seed + 100).standard_normal(ent.shape)
    ent[..., -1] = ent[..., 0]
    qg.set_forcing(ent, wekp, txis, txin)
    print(f"[diag] continuity {w.name}", flush=True)
    qg.run(w.tdt, 30, 0, w.supp_every)
    qg.sync()
    d = qg.diagnose(w.dt)
    _, _, p, _ = qg.get_state()
    nl = w.nl
    int_p = [M.int_P_dA(p[k], w.dx, w.dx) for k in range(nl)]
    dpi = qg.get_constr()[: nl - 1]
    assert np.any(dpi != 0.0)   # the entrainment moved the interfaces
    xl, yl = (w.nxp - 1) * w.dx, (w.nyp - 1) * w.dx
    ermas, emfr = M.check_continuity(int_p, dpi, w.gp, w.params()["dz_sign"], xl, yl, w.tdt)
    for k in range(nl - 1):
        esum = abs(w.params()["dz_sign"] * (int_p[k] - int_p[k + 1])) + abs(dpi[k])
        assert abs(d["ermas"][k] - ermas[k]) <= 1e-12 * esum, (k, d["ermas"], ermas)
        assert abs(d["emfr"][k] - emfr[k]) <= 4e-12 + 1e-9 * abs(emfr[k]), (k, d["emfr"], emfr)
    print(f"[diag] ermas {d['ermas']} emfr {d['emfr']}", flush=True)


@pytest.mark.parametrize("make,steps", [(W.small_box, 12), (W.small_channel, 12), (W.c2_box513, 3)],
                         ids=["box33", "chan33x17", "box513"])
def test_energetics_match_oracle(gpu, make, steps):
    """oqg_diagnose_energy against the numpy restatement of qg_diagno's
    energetics (oracle/monitor.py) on a state evolved with lateral friction,
    wind stress set on the device and seeded interface entrainment: the
    oracle reads the device's own resident forcing and state."""
    from dataclasses import replace

    w = make()
    w = replace(w, ah2=[150.0] * w.nl)
    qg = w.make_qg()
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    ent, wekp, txis, txin = w.forcing()
    rng = np.random.default_rng(w.seed + 7)
    ent = 1e-6 * rng.standard_normal(ent.shape)
    if w.cyclic:
        ent[..., -1] = ent[..., 0]
    qg.set_forcing(ent, wekp, txis, txin)
    tx = 1e-4 * (1.0 + rng.standard_normal((w.nyp, w.nxp)))
    ty = 2e-5 * rng.standard_normal((w.nyp, w.nxp))
    if w.cyclic:
        tx[:, -1], ty[:, -1] = tx[:, 0], ty[:, 0]
    qg.set_wind(tx, ty)
    print(f"[energy] {w.name} after {steps} steps", flush=True)
    qg.run(w.tdt, steps, 0, w.supp_every)
    qg.sync()
    e = qg.diagnose_energy()
    _, _, p, pm = qg.get_state()
    wk, _, _ = qg.get_forcing()
    gx, gy = qg.get_wind()
    ent1 = qg.get_ent()
    dz = w.params()["dz_sign"]
    kt = 0 if dz == 1 else w.nl - 1
    o = M.energetics(p, pm, wk, ent1, gx, gy, w.dx, w.dx, w.fnot, qg.bcco, w.cyclic, w.h, w.gp, qg.ah2, qg.ah4,
                     qg.delek, kt, dz, qg.rho)
    for name in ("ah2d", "ah4d", "val", "btdg", "utau", "wepm", "wapm", "entm", "enam", "pken"):
        a, b = np.atleast_1d(e[name]), np.atleast_1d(o[name])
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-9 * scale, (name, a, b)
        if not (name == "btdg" and qg.delek == 0.0):   # the channel has no bottom Ekman layer
            assert np.all(b != 0.0), name   # every quantity is exercised
    assert np.array_equal(e["pos"], o["pos"]), (e["pos"], o["pos"])
    print(f"[energy] ah4d {e['ah4d']} btdg {e['btdg']:.6e} utau {e['utau']:.6e}", flush=True)
