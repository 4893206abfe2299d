"""The Fortran host path: openqg_amd/fortran/test_oqg_c.F90 drives liboqg.so
through the ISO_C_BINDING module oqg_c.  Its result must equal, bitwise, the
same case run through the Python binding, and match the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_err
from oracle import cpuref as cr
from openqg_amd import build as B
from openqg_amd.qg import QG, Box

pytestmark = pytest.mark.gpu

N, NL = 65, 3


def _python_run(p0, wekp):
    b = Box(N, N, NL, 30.0e3, 30.0e3, 1.0e-4, 2.0e-11, [300.0, 1100.0, 2600.0], cyclic=False, dz_sign=-1)
    qg = QG(b, 0.5, 2.0, [0.0] * NL, [2.0e10] * NL, [0.05, 0.025], tau_sign=1)
    qg.set_state(p=p0, pm=p0)
    qg.init_state()
    ent = np.zeros((NL - 1, N, N))
    qg.set_forcing(ent, wekp, 0.0, 0.0)
    for nt in range(1, 11):
        qg.step(7200.0)
        if nt == 1:
            qg.suppress()
    out = qg.get_state()
    qg.close()
    return out


def test_fortran_program_matches_python_binding_and_oracle(gpu, tmp_path):
    exe = B.FORTRAN_TEST
    if not os.path.exists(exe):
        pytest.skip("Fortran test program not built (amdflang absent when build() ran)")
    out = tmp_path / "fortran_state.bin"
    print(f"[fortran] running {exe}", flush=True)
    r = subprocess.run([exe, str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    print(r.stdout.decode(errors="replace"), flush=True)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    f = N * N * NL
    p0 = raw[:f].reshape(NL, N, N)
    wekp = raw[f:f + N * N].reshape(N, N)
    fq, fqm, fp, fpm = (raw[f + N * N + i * f: f + N * N + (i + 1) * f].reshape(NL, N, N) for i in range(4))

    gq, gqm, gp, gpm = _python_run(p0, wekp)
    for a, b in ((fq, gq), (fqm, gqm), (fp, gp), (fpm, gpm)):
        assert np.array_equal(a, b)

    prm = dict(nxp=N, nyp=N, nl=NL, cyclic=0, dx=30.0e3, dy=30.0e3, fnot=1.0e-4, beta=2.0e-11,
               h=[300.0, 1100.0, 2600.0], gp=[0.05, 0.025], ah2=[0.0] * NL, ah4=[2.0e10] * NL,
               bcco=0.5, delek=2.0, dz_sign=-1, tau_sign=1, k_topo=NL)
    o = cr.Basin(prm)
    o.set_state(p=p0, pm=p0)
    o.init_state()
    ent = np.zeros((NL - 1, N, N))
    for nt in range(1, 11):
        o.step(7200.0, ent, wekp, 0.0, 0.0)
        if nt == 1:
            o.suppress()
    oq, oqm, op, opm = o.get_state()
    assert rel_err(fq, oq) <= 1e-10
    assert rel_err(fp, op) <= 1e-10
