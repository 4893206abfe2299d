"""The ocean mixed layer on the device (oqg_oml_*, oqg_run_ocean) against the
oracle's C restatement of step_oml (omlsubs.F90:69-119) with adv_diff
(mixed_layer.F90:16-95) and the Ekman velocities (ekman.F90:73-80).

Bar (FP64): relative max-norm error <= 1e-12 on sst and the entrainment after one
step from an identical state, on boxes and channels, with and without flux walls,
at 33^2 and at the 513^2 ocean's full size; <= 1e-10 over a 40-step coupled run
(mixed layer + basin, with suppression).  The graph-replayed ocean run equals the
same steps launched one by one, bit for bit.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle import cpuref as cr
from openqg_amd import workloads as W
from openqg_amd.mixed import OceanMixedLayer, init_temp

pytestmark = pytest.mark.gpu


def _stress(w, rng, amp=1e-4):
    tx, ty = (amp * rng.standard_normal((w.nyp, w.nxp)) for _ in range(2))
    tx += amp * np.cos(np.pi * np.arange(w.nyp) / (w.nyp - 1))[:, None]   # a gyre-like zonal part
    if w.cyclic:
        tx[:, -1] = tx[:, 0]
        ty[:, -1] = ty[:, 0]
    return tx, ty


def _setup(w, seed, fluxn=False, fluxs=True, hm=100.0):
    """A basin from workload `w`, its mixed layer (rad-balance-like profile: T1_rel = -2 K,
    T2_rel = -12 K, dT_NS = 8 K) and seeded sst, sstm, fnet and wind stress."""
    rng = np.random.default_rng(seed)
    qg = w.make_qg()
    b = qg.b
    temp = init_temp(b, -2.0, -12.0, 8.0, 1.0e3, 4.0e3)
    oml = OceanMixedLayer(qg, temp, fluxn=fluxn and b.fnot < 0, fluxs=fluxs and b.fnot > 0, st2d=380.0, st4d=4.0e10,
                          ycexp=1.0, hm=hm)
    nyt, nxt = w.nyp - 1, w.nxp - 1
    yt = np.linspace(-1.0, 1.0, nyt)[:, None]
    sst = -4.0 * yt + 0.5 * rng.standard_normal((nyt, nxt))
    sstm = sst + 0.05 * rng.standard_normal((nyt, nxt))
    # a heat flux that cools some points below T1_rel within a step (convection)
    fnet = -40.0 * yt + 30.0 * rng.standard_normal((nyt, nxt))
    fnet[rng.random((nyt, nxt)) < 0.05] = -2.0e5
    tx, ty = _stress(w, rng)
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    qg.set_wind(tx, ty)
    oml.set_state(sst, sstm)
    oml.set_fnet(fnet)
    return qg, oml, dict(sst=sst, sstm=sstm, fnet=fnet, tx=tx, ty=ty)


CASES = [(W.small_box, False, True), (W.small_box, False, False), (W.small_channel, False, False),
         (W.small_channel, False, True), (W.c2_box513, False, True)]


@pytest.mark.parametrize("make,fluxn,fluxs", CASES,
                         ids=["box33-fluxs", "box33-walls", "chan33x17", "chan33x17-fluxs", "box513-fluxs"])
def test_oml_step_matches_oracle(gpu, make, fluxn, fluxs):
    w = make()
    qg, oml, f = _setup(w, 11, fluxn, fluxs)
    print(f"[oml] {w.name} fluxn={fluxn} fluxs={fluxs}", flush=True)
    cfg = w.params()
    p1 = qg.get_state()[2][0]
    tdto = w.tdt
    es, esm, ee = cr.oml_step(cfg, oml.params(), p1, f["tx"], f["ty"], f["fnet"], tdto, f["sst"], f["sstm"])
    oml.step(tdto)
    gs, gsm = oml.get_state()
    ge = qg.get_ent()
    np.testing.assert_array_equal(gsm, f["sst"])          # sstm <- sst is a copy
    assert rel_err(gs, es) <= 1e-12
    assert rel_err(ge, ee) <= 1e-12
    assert np.any(gs == oml.temp.T1_rel), "the case should convect somewhere"
    # without the entrainment, the resident field is left as it was
    oml.step(tdto, compute_ent=False)
    np.testing.assert_array_equal(qg.get_ent(), ge)


def test_ocean_run_graph_equals_eager(gpu):
    """oqg_run_ocean replays captured graphs of (step_oml, step_qg, suppress_ocean);
    the same steps issued one call at a time give the same bits."""
    w = W.small_box(65, 3)
    runs = []
    for mode in ("graph", "eager"):
        qg, oml, _ = _setup(w, 5)
        print(f"[oml-run] {w.name} {mode}", flush=True)
        n, supp = 60, 25
        if mode == "graph":
            oml.run(w.tdt, n, 0, supp)
        else:
            for nt in range(1, n + 1):
                oml.step(w.tdt)
                qg.step(w.tdt)
                if (nt - 1) % supp == 0:
                    qg.suppress()
                    oml.suppress()
        runs.append((oml.get_state(), qg.get_state(), qg.get_ent()))
        oml.close()
        qg.close()
    (s0, q0, e0), (s1, q1, e1) = runs
    for a, b in zip(s0 + q0 + (e0,), s1 + q1 + (e1,)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("make", [W.small_box, W.small_channel], ids=["box33", "chan33x17"])
def test_coupled_ocean_run_matches_oracle(gpu, make):
    """40 steps of step_ocn + suppress_ocean every 25 steps on the device against the
    oracle's step_oml feeding the oracle's step_qg (driver.F90:89-110, 160-171)."""
    w = make()
    qg, oml, f = _setup(w, 17, fluxs=not w.cyclic)
    print(f"[oml-coupled] {w.name}", flush=True)
    cfg = w.params()
    o = cr.Basin(cfg)
    o.set_state(*qg.get_state())
    o.set_constr(qg.get_constr())
    wekp, txis, txin = cr.ekman_wekp(cfg, f["tx"], f["ty"])
    sst, sstm = f["sst"], f["sstm"]
    n, supp = 40, 25
    oml.run(w.tdt, n, 0, supp)
    for nt in range(1, n + 1):
        p1 = o.get_state()[2][0]
        sst, sstm, e1 = cr.oml_step(cfg, oml.params(), p1, f["tx"], f["ty"], f["fnet"], w.tdt, sst, sstm)
        ent = np.zeros((w.nl - 1, w.nyp, w.nxp))
        ent[0] = e1
        o.step(w.tdt, ent, wekp, txis, txin)
        if (nt - 1) % supp == 0:
            o.suppress()
            sst = 0.5 * (sst + sstm)
    gs, gsm = oml.get_state()
    gq, _, gp, _ = qg.get_state()
    oq, _, op, _ = o.get_state()
    assert rel_err(gs, sst) <= 1e-10 and rel_err(gsm, sstm) <= 1e-10
    assert rel_err(gq, oq) <= 1e-10 and rel_err(gp, op) <= 1e-10
    assert rel_err(qg.get_ent(), e1) <= 1e-10


def test_oml_rejects_bad_arguments(gpu):
    w = W.small_box()
    qg = w.make_qg()
    temp = init_temp(qg.b, -2.0, -12.0, 8.0, 1.0e3, 4.0e3)
    with pytest.raises(ValueError):
        OceanMixedLayer(qg, temp, fluxn=True, fluxs=False, st2d=0.0, st4d=0.0, ycexp=1.0, hm=100.0)
    with pytest.raises(RuntimeError):
        OceanMixedLayer(qg, temp, fluxn=False, fluxs=False, st2d=0.0, st4d=0.0, ycexp=1.0, hm=0.0)
    oml = OceanMixedLayer(qg, temp, fluxn=False, fluxs=False, st2d=0.0, st4d=0.0, ycexp=1.0, hm=50.0)
    with pytest.raises(ValueError):
        oml.set_state(np.zeros((3, 3)))
    qg.close()   # closes the bound mixed layer first
    assert oml._h is None
