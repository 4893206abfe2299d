"""The batched gradient (openqg_amd.adjoint.gradient_batch): members run
concurrently from host threads, each on its own context and stream, must give
exactly the gradients of the members run one at a time."""
import dataclasses

import numpy as np
import pytest

from openqg_amd import workloads as W
from openqg_amd.adjoint import gradient, gradient_batch
from openqg_amd.forcing import double_gyre_tau

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("snapshots", [None, 3], ids=["store_all", "revolve3"])
def test_batched_gradients_equal_single_runs(gpu, snapshots):
    base = W.small_box()
    members = [dataclasses.replace(base, seed=base.seed + i, tau0=base.tau0 * (1.0 + 0.1 * i)) for i in range(4)]
    print(f"[adjoint batch] {base.name} x{len(members)} snapshots={snapshots}", flush=True)
    qgs, q0s, txs, tys = [], [], [], []
    for w in members:
        qg = w.make_qg()
        p = w.initial_p()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        q0s.append(qg.get_state()[0])
        tx, ty = double_gyre_tau(w.nyp, w.nxp, w.dx, w.tau0)
        txs.append(tx)
        tys.append(ty)
        qgs.append(qg)
    n = 40
    ref = []
    for i, w in enumerate(members):
        ref.append(gradient(qgs[i], q0s[i], txs[i], tys[i], w.tdt, n, w.supp_every, snapshots))
    for rep in range(2):
        got = gradient_batch(qgs, q0s, txs, tys, base.tdt, n, base.supp_every, snapshots)
        for i, (a, b) in enumerate(zip(got, ref)):
            assert a[0] == b[0], f"member {i} J"
            for k in (1, 2, 3):
                assert np.array_equal(a[k], b[k]), f"member {i} output {k}"
            assert a[4] == b[4]
    assert len({r[0] for r in ref}) == len(members)   # the members differ
    with pytest.raises(ValueError):
        gradient_batch([qgs[0], qgs[0]], q0s[:2], txs[:2], tys[:2], base.tdt, n)
    for qg in qgs:
        qg.close()
