"""The drop-in Fortran `module core` (openqg_amd/fortran/core.F90) compiled against a
test-only stand-in for OpenQG's `module qg` (tests/fortran/qg_types.F90) and driven
by tests/fortran/test_core.F90 the way driver.F90 drives step_qg: a basin filled
with init_qg's precompute (here the oracle's), step_qg every step, suppress_comp_mode
every `supp` steps.

Checked, for a box and a channel, in both of core.F90's modes (host-authoritative by
default, device-authoritative with OQG_RESIDENT=1):
  - the two modes give the same bits;
  - they equal, bitwise, the same run through the Python binding built from the same
    precompute;
  - they match the oracle run from the same inputs (relative max-norm <= 1e-10 after
    10 steps, the trajectory bar of the other multi-step parity tests).

Input file layout (native-endian records, in this order): int32 nxp, nyp, nl, cyclic,
nsteps, supp, k_topo, tau_sign, dz_sign; float64 dx, dy, fnot, beta, bcco, delek, tdt,
txis, txin; h(nl), gp(nl-1), ah2(nl), ah4(nl); ddyn(nxp,nyp); amat, ctl2m, ctm2l (nl,nl)
and rdm2(nl); for a box hom_sol_bc(nxp,nyp,nl-1), for a channel hom_sol_bt(nyp),
hom_sol_bc1 and hom_sol_bc2 (nyp,nl-1) and hc1s, hc2s, hc1n, hc2n per mode then hbsi;
q, qm, p, pm (nxp,nyp,nl); dpi, dpip (nl-1); for a channel cs, cn, csp, cnp (nl);
ent (nxp,nyp,nl); wekp (nxp,nyp).  All arrays in Fortran order.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_err
from oracle import cpuref as cr
from openqg_amd import build as B
from openqg_amd import workloads as W

pytestmark = pytest.mark.gpu

NSTEPS, SUPP = 10, 4


def _case(make):
    w = make()
    o = cr.Basin(w.params())
    p = w.initial_p()
    o.set_state(p=p, pm=p)
    o.init_state()
    ent, wekp, txis, txin = w.forcing()
    ent = 2e-6 * np.random.default_rng(w.seed + 50).standard_normal(ent.shape)
    if w.cyclic:
        ent[..., -1] = ent[..., 0]
    return w, o, ent, wekp, txis, txin


def _write_input(path, w, o, ent, wekp, txis, txin):
    nl = w.nl
    k_topo = nl if w.dz_sign == -1 else 1
    amat, ctl2m, ctm2l, rdm2 = o.modes()   # [row][col] = (i, j)
    with open(path, "wb") as f:
        np.array([w.nxp, w.nyp, nl, int(w.cyclic), NSTEPS, SUPP, k_topo, w.tau_sign, w.dz_sign],
                 dtype=np.int32).tofile(f)
        np.array([w.dx, w.dx, w.fnot, w.beta, w.bcco, w.delek, w.tdt, txis, txin]).tofile(f)
        for a in (w.h, w.gp, w.ah2, w.ah4):
            np.asarray(a, dtype=np.float64).tofile(f)
        np.zeros((w.nyp, w.nxp)).tofile(f)                       # flat: ddyn = 0
        for a in (amat, ctl2m, ctm2l):
            np.ascontiguousarray(a.T).tofile(f)                  # column-major (nl,nl)
        rdm2.tofile(f)
        hom = o.homog()
        if w.cyclic:
            for a in hom:
                np.ascontiguousarray(a).tofile(f)
        else:
            np.ascontiguousarray(hom).tofile(f)
        for a in o.get_state():
            np.ascontiguousarray(a).tofile(f)
        c = o.get_constr()
        c[: 2 * (nl - 1)].tofile(f)
        if w.cyclic:
            c[2 * (nl - 1): 2 * (nl - 1) + 4 * nl].tofile(f)
        full = np.zeros((nl, w.nyp, w.nxp))
        full[: nl - 1] = ent
        full.tofile(f)
        np.ascontiguousarray(wekp).tofile(f)


def _read_output(path, w):
    raw = np.fromfile(path, dtype=np.float64)
    nl, f = w.nl, w.nl * w.nyp * w.nxp
    fields = [raw[i * f:(i + 1) * f].reshape(nl, w.nyp, w.nxp) for i in range(4)]
    rest = raw[4 * f:]
    assert rest.size == 2 * (nl - 1) + (4 * nl if w.cyclic else 0)
    return fields, rest


def _python_run(w, o, ent, wekp, txis, txin):
    """The same run through the Python binding, built from the same precompute."""
    amat, ctl2m, ctm2l, rdm2 = o.modes()
    pre = dict(amat=amat, ctl2m=ctl2m, ctm2l=ctm2l, rdm2=rdm2)
    hom = o.homog()
    if w.cyclic:
        pre.update(hom_bt=hom[0], hom_bc1=hom[1], hom_bc2=hom[2], hom_coef=hom[3])
    else:
        pre["hom_sol_bc"] = hom
    qg = w.make_qg(precompute=pre)
    qg.set_state(*o.get_state())
    qg.set_constr(o.get_constr())
    qg.set_forcing(ent, wekp, txis, txin)
    for nt in range(1, NSTEPS + 1):
        qg.step(w.tdt)
        if (nt - 1) % SUPP == 0:
            qg.suppress()
    out = qg.get_state(), qg.get_constr()
    qg.close()
    return out


@pytest.mark.parametrize("make", [W.small_box, W.small_channel], ids=["box33", "chan33x17"])
def test_core_f90_drop_in(gpu, tmp_path, make):
    exe = B.FORTRAN_CORE_TEST
    if not os.path.exists(exe):
        pytest.skip("core.F90 test program not built (amdflang absent when build() ran)")
    w, o, ent, wekp, txis, txin = _case(make)
    inp = tmp_path / "in.bin"
    _write_input(inp, w, o, ent, wekp, txis, txin)
    runs = {}
    for mode in ("default", "resident"):
        env = dict(os.environ)
        env.pop("OQG_RESIDENT", None)
        if mode == "resident":
            env["OQG_RESIDENT"] = "1"
        out = tmp_path / f"out_{mode}.bin"
        print(f"[core.F90] {w.name} {mode}", flush=True)
        r = subprocess.run([exe, str(inp), str(out)], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           timeout=120)
        print(r.stdout.decode(errors="replace"), flush=True)
        assert r.returncode == 0
        runs[mode] = _read_output(out, w)
    (fd, cd), (fr, cr_) = runs["default"], runs["resident"]
    for a, b in zip(fd, fr):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(cd, cr_)
    # the Python binding, same precompute, same steps: same bits
    (pq, pqm, pp, ppm), pc = _python_run(w, o, ent, wekp, txis, txin)
    for a, b in zip(fd, (pq, pqm, pp, ppm)):
        np.testing.assert_array_equal(a, b)
    nl = w.nl
    np.testing.assert_array_equal(cd[: 2 * (nl - 1)], pc[: 2 * (nl - 1)])
    if w.cyclic:
        np.testing.assert_array_equal(cd[2 * (nl - 1):], pc[2 * (nl - 1): 2 * (nl - 1) + 4 * nl])
    # the oracle from the same inputs
    for nt in range(1, NSTEPS + 1):
        o.step(w.tdt, ent, wekp, txis, txin)
        if (nt - 1) % SUPP == 0:
            o.suppress()
    oq, oqm, op, opm = o.get_state()
    assert rel_err(fd[0], oq) <= 1e-10 and rel_err(fd[2], op) <= 1e-10
    assert rel_err(fd[1], oqm) <= 1e-10 and rel_err(fd[3], opm) <= 1e-10
    oc = o.get_constr()
    scale = np.abs(oc).max()
    assert np.abs(cd[: 2 * (nl - 1)] - oc[: 2 * (nl - 1)]).max() <= 1e-9 * scale
