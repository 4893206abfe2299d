"""Restart through a file (src/state.F90:29-44,82-112; src/openqg.F90:237-243).

A basin runs some steps on the GPU, its p and pm go to a restart file, and a
fresh context loads them and derives q, qm and the constraint scalars on the
GPU.  The loaded pressures must equal the saved ones bit for bit, the derived
state must match the oracle's init_pv/init_constraint/init_mass_constr of the
same p and pm to 1e-12, and the steps after the restart must keep the 1e-12
per-step bar against the oracle.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle import cpuref as cr
from openqg_amd import workloads as W
from openqg_amd.state import restart_qg, save_qg

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _constr_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b), np.abs(b).max() * 1e-3 + 1e-300)
    return float(np.max(np.abs(a - b) / scale))


@pytest.mark.parametrize("make", [W.small_box, W.small_channel], ids=["box33", "chan33x17"])
def test_restart_file_resumes_the_run(gpu, tmp_path, make):
    w = make()
    print(f"[restart] {w.name}", flush=True)
    ent, wekp, txis, txin = w.forcing()
    a = w.make_qg()
    p0 = w.initial_p()
    a.set_state(p=p0, pm=p0)
    a.init_state()
    a.set_forcing(ent, wekp, txis, txin)
    a.run(w.tdt, 12, 0, w.supp_every)
    a.sync()
    _, _, pa, pma = a.get_state()
    name = f"{w.name}_qg_restart.nc"
    save_qg(str(tmp_path), name, a)

    b = w.make_qg()
    restart_qg(str(tmp_path), name, b)
    b.set_forcing(ent, wekp, txis, txin)
    q, qm, p, pm = b.get_state()
    assert p.tobytes() == pa.tobytes() and pm.tobytes() == pma.tobytes()

    o = cr.Basin(w.params())
    o.set_state(p=pa, pm=pma)
    o.init_state()
    oq, oqm, op, opm = o.get_state()
    assert rel_err(q, oq) <= TOL and rel_err(qm, oqm) <= TOL
    assert _constr_err(b.get_constr(), o.get_constr()) <= 1e-10

    for nt in range(1, 6):
        gq, gqm, gp, gpm = b.get_state()
        o.set_state(gq, gqm, gp, gpm)
        o.set_constr(b.get_constr())
        b.step(w.tdt)
        o.step(w.tdt, ent, wekp, txis, txin)
        gq, _, gp, _ = b.get_state()
        oq, _, op, _ = o.get_state()
        assert rel_err(gq, oq) <= TOL, f"step {nt} after restart: q"
        assert rel_err(gp, op) <= TOL, f"step {nt} after restart: p"
