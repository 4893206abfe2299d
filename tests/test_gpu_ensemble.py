"""Batched ensembles (oqg_ens_*): every member must follow the trajectory it
follows when stepped alone through the plain path.  With the default launch
schedule the kernels and their arguments are the same, so the results are
bitwise equal; the single-y-solve schedule partitions the short columns of the
small members differently and agrees to round-off."""
import numpy as np
import pytest

from conftest import rel_err
from openqg_amd import workloads as W
from openqg_amd.ensemble import ENS_LARGEST_FIRST, ENS_ONE_YSOLVE, Ensemble

pytestmark = pytest.mark.gpu


def _members():
    ws = [W.small_box(n=33, nl=2, seed=21), W.small_box(n=65, nl=3, seed=22), W.small_box(n=97, nl=4, seed=23),
          W.small_box(n=129, nl=3, seed=24), W.small_channel(33, 17, 3, seed=25), W.small_box(n=65, nl=4, seed=26)]
    ws[1].ah4 = [5.0e10] * 3
    ws[3].tau0 = 0.15
    return ws


def _make(ws):
    qgs = []
    for w in ws:
        qg = w.make_qg()
        p = w.initial_p()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        ent, wekp, txis, txin = w.forcing()
        qg.set_forcing(ent, wekp, txis, txin)
        qgs.append(qg)
    return qgs


def _alone(ws, nsteps, supp):
    out = []
    for w, qg, n in zip(ws, _make(ws), nsteps):
        for nt in range(1, n + 1):
            qg.step(w.tdt)
            if supp and (nt - 1) % supp == 0:
                qg.suppress()
        out.append((qg.get_state(), qg.get_constr()))
        qg.close()
    return out


@pytest.mark.parametrize("schedule", [0, ENS_LARGEST_FIRST, ENS_ONE_YSOLVE, ENS_ONE_YSOLVE | ENS_LARGEST_FIRST])
def test_ensemble_equals_members_stepped_alone(gpu, schedule):
    ws = _members()
    nsteps = [7, 12, 12, 5, 9, 12]    # uneven: members drop out of the batch as they finish
    supp = 4
    ref = _alone(ws, nsteps, supp)
    qgs = _make(ws)
    ens = Ensemble(qgs)
    ens.schedule = schedule
    print(f"[ensemble] schedule {schedule}", flush=True)
    ens.run([w.tdt for w in ws], nsteps, 0, supp)
    ens.sync()
    for i, (qg, (st, con)) in enumerate(zip(qgs, ref)):
        got = qg.get_state()
        for name, a, b in zip(("q", "qm", "p", "pm"), got, st):
            if schedule & ENS_ONE_YSOLVE:
                assert rel_err(a, b) <= 1e-12, f"member {i} {name}"
            else:
                assert np.array_equal(a, b), f"member {i} {name}"
        assert np.allclose(qg.get_constr(), con, rtol=1e-10, atol=0.0)
    ens.close()


def test_ensemble_runs_continue_and_tune_restores_state(gpu):
    ws = _members()[:4]
    ref = _alone(ws, [10] * 4, 3)
    qgs = _make(ws)
    ens = Ensemble(qgs)
    tdt = [w.tdt for w in ws]
    ens.run(tdt, 4, 0, 3)         # two runs with an odd split point = one run of 10 steps
    ens.sync()
    before = [qg.get_state() for qg in qgs]
    best, ms = ens.tune(tdt, trial_steps=4, supp_every=3)
    assert ens.schedule == best and len(ms) == 16 and all(t > 0 for t in ms)
    assert ens.tile_option in (0, 4, 8, 16)
    if ens.tile_option:
        assert all(qg.tile_rows == ens.tile_option for qg in qgs)
    for qg, b in zip(qgs, before):
        for x, y in zip(qg.get_state(), b):
            assert np.array_equal(x, y)
    ens.schedule = 0
    ens.run(tdt, 6, 4, 3)
    ens.sync()
    for qg, (st, _) in zip(qgs, ref):
        for a, b in zip(qg.get_state(), st):
            assert np.array_equal(a, b)
    ens.close()


@pytest.mark.parametrize("solo", [False, True], ids=["groups", "groups_and_solo"])
def test_ensemble_groups_equal_members_stepped_alone(gpu, solo):
    from openqg_amd.ensemble import EnsembleGroups

    ws = _members()
    nsteps = [7, 12, 12, 5, 9, 12]
    ref = _alone(ws, nsteps, 4)
    qgs = _make(ws)
    big = max(w.nxp * w.nyp for w in ws)
    grp = EnsembleGroups(qgs, 2, solo_points=big if solo else None)
    assert sorted(grp.solo + [i for g in grp.index for i in g]) == list(range(len(ws)))
    assert len(grp.index) == 2 and (len(grp.solo) > 0) == solo
    print(f"[ensemble] groups {grp.index} solo {grp.solo}", flush=True)
    grp.run([w.tdt for w in ws], nsteps, 0, 4)
    grp.sync()
    for i, (qg, (st, con)) in enumerate(zip(qgs, ref)):
        for name, a, b in zip(("q", "qm", "p", "pm"), qg.get_state(), st):
            assert np.array_equal(a, b), f"member {i} {name}"
        assert np.allclose(qg.get_constr(), con, rtol=1e-10, atol=0.0)
    grp.close()


def test_member_read_after_run_without_sync(gpu):
    """A member read on its own stream right after Ensemble.run (no sync) sees
    the run's result: the run ends by making every member's stream wait for
    the ensemble's."""
    ws = _members()
    nsteps = [20, 20, 20, 20, 20, 20]
    ref = _alone(ws, nsteps, 25)
    qgs = _make(ws)
    ens = Ensemble(qgs)
    ens.run([w.tdt for w in ws], nsteps, 0, 25)
    for i, (qg, (st, _)) in enumerate(zip(qgs, ref)):
        got = qg.get_state()
        for name, a, b in zip(("q", "qm", "p", "pm"), got, st):
            assert np.array_equal(a, b), f"member {i} {name}"
    ens.close()
