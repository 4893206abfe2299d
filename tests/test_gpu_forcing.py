"""The wind-stress forcing producer on the device (oqg_set_wind, oqg_adj_wind)
against the oracle's C restatement of ekman_from_tau + map_T_to_P and the
host numpy one, for boxes and channels, and its transpose by a dot-product
test and against the numpy transpose."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import cpuref as cr
from openqg_amd import workloads as W
from openqg_amd.adjoint import AdjointState
from openqg_amd.forcing import ekman_wekp, ekman_wekp_adjoint

pytestmark = pytest.mark.gpu


def _stress(w, seed):
    rng = np.random.default_rng(seed)
    tx, ty = (1e-4 * rng.standard_normal((w.nyp, w.nxp)) for _ in range(2))
    if w.cyclic:   # the duplicate column nxp equals column 1
        tx[:, -1] = tx[:, 0]
        ty[:, -1] = ty[:, 0]
    return tx, ty


@pytest.mark.parametrize("make", [W.small_box, W.small_channel, W.c2_box513, W.atmos_channel],
                         ids=["box33", "chan33x17", "box513", "atmos97x49"])
def test_wind_forcing_matches_oracle(gpu, make):
    w = make()
    qg = w.make_qg()
    tx, ty = _stress(w, 21)
    print(f"[forcing] {w.name}", flush=True)
    qg.set_wind(tx, ty)
    wd, txis, txin = qg.get_forcing()
    wo, txis_o, txin_o = cr.ekman_wekp(w.params(), tx, ty)
    wn, txis_n, txin_n = ekman_wekp(w.nxp, w.nyp, w.dx, w.dx, w.fnot, w.cyclic, tx, ty)
    # same operations in the same order as the oracle: bitwise on every point
    np.testing.assert_array_equal(wd, wo)
    assert rel_err(wd, wn) <= 1e-15
    if w.cyclic:
        scale = max(abs(txis_o), abs(txin_o))
        assert abs(txis - txis_o) <= 1e-14 * scale and abs(txin - txin_o) <= 1e-14 * scale
        assert abs(txis - txis_n) <= 1e-14 * scale
    # the resident entrainment is untouched, and a step sees the new forcing
    ent = np.zeros((max(w.nl - 1, 1), w.nyp, w.nxp))
    p = w.initial_p()
    for b in (qg,):
        b.set_state(p=p, pm=p)
        b.init_state()
    o = cr.Basin(w.params())
    o.set_state(*qg.get_state())
    o.set_constr(qg.get_constr())
    qg.set_forcing(ent, None, txis_o, txin_o)
    qg.set_wind(tx, ty)
    qg.step(w.tdt)
    o.step(w.tdt, ent, wo, txis_o, txin_o)
    gq, _, gp, _ = qg.get_state()
    oq, _, op, _ = o.get_state()
    assert rel_err(gq, oq) <= 1e-12 and rel_err(gp, op) <= 1e-12


def test_wind_adjoint_is_the_transpose(gpu):
    w = W.small_box(33, 3)
    qg = w.make_qg()
    adj = AdjointState(qg)
    rng = np.random.default_rng(3)
    lw = rng.standard_normal((w.nyp, w.nxp))
    # put lw in the adjoint of wekp: one adjoint step from a zero adjoint state
    # with only the forcing's contribution is not needed; set it directly
    # through the identity <lw, wekp(tau)> = <adj(lw), tau>
    tx, ty = _stress(w, 4)
    qg.set_wind(tx, ty)
    wd, _, _ = qg.get_forcing()
    ax, ay = ekman_wekp_adjoint(w.nxp, w.nyp, w.dx, w.dx, w.fnot, lw)
    lhs = float(np.sum(lw * wd))
    rhs = float(np.sum(ax * tx) + np.sum(ay * ty))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # the device transpose of the accumulated adjoint of wekp after one adjoint
    # step equals the numpy transpose of the same field
    qg.set_state(p=w.initial_p(), pm=w.initial_p())
    qg.init_state()
    lq = rng.standard_normal(qg.shape)
    adj.alloc(1)
    adj.save(0)
    qg.step(w.tdt)
    adj.set(lq=lq)
    adj.zero_wekp()
    adj.load(0)
    adj.step(w.tdt, False)
    lwek = adj.wekp()
    gx, gy = adj.wind()
    nx_, ny_ = ekman_wekp_adjoint(w.nxp, w.nyp, w.dx, w.dx, w.fnot, lwek)
    assert np.abs(lwek).max() > 0
    assert rel_err(gx, nx_) <= 1e-13 and rel_err(gy, ny_) <= 1e-13
