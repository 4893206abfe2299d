"""The coupled model's set-up radiation (openqg_amd.radiation: load_rad, radiat,
radsubs.F90:63-676; init_entrain_coeff, amlsubs.F90:93-129) with the double-gyre
configuration's values (data/basin/atm_dg.cdl, ocn_dg.cdl): the mean temperatures close
their balances, the radiative-balance multipliers solve their linear system, and the
perturbation coefficients are the derivatives of the mean fluxes they linearise."""
import math

import numpy as np

from openqg_amd import radiation as R
from openqg_amd.qg import Box


def _ga(nl=3, hm=1000.0):
    return Box(97, 49, nl, 120e3, 120e3, 1e-4, 2e-11, [2000.0, 3000.0, 4000.0][:nl], cyclic=True, dz_sign=1, hm=hm)


def _rad(ga):
    return R.load_rad(ga, -210.0, 80.0, [200.0, 3.0e4, 3.0e4, 4.0e4], 1.0e-2, 35.0, [330.0, 340.0, 350.0])


def test_mean_state_closes_the_balances():
    ga = _ga()
    rad = _rad(ga)
    rbtmat, rbtmoc, toc, tat = R.radiat(ga, rad)
    # the ocean mixed layer: sensible/latent + its emission = the atmosphere's downward + solar
    lhs = rad.xlamda * (toc - tat) + R.STEFAN * toc ** 4
    rhs = R.SIGOV2 * tat ** 4 - rad.fsbar
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)
    assert 250.0 < toc < 350.0 and 250.0 < tat < 350.0
    assert rad.fspco == math.copysign(80.0, ga.fnot)
    # D0up and Dmdown linearise the ocean's and the mixed layer's black-body fluxes
    assert np.isclose(rad.D0up, 4 * R.STEFAN * toc ** 3, rtol=1e-14)
    assert np.isclose(rad.Dmdown, -2 * R.STEFAN * tat ** 3, rtol=1e-14)
    # the transmissions shrink the mixed layer's upward sensitivity layer by layer
    tauk = np.exp(-np.array([ga.hm, ga.h[0] - ga.hm, ga.h[1], ga.h[2]]) / rad.zopt)
    assert np.allclose(rad.Dup[1:] / rad.Dup[:-1], tauk[1:], rtol=1e-14)


def test_radiative_balance_multipliers_solve_their_system():
    ga = _ga()
    rad = _rad(ga)
    rbtmat, rbtmoc, _, _ = R.radiat(ga, rad)
    nl = ga.nl
    f = np.concatenate([rad.rbetat, [rbtmat]])
    A = np.zeros((nl, nl))
    A[0, : nl - 1] = [R.Adown(rad, 1, c) for c in range(1, nl)]
    A[0, nl - 1] = rad.Dup[0]
    for k in range(2, nl):
        A[k - 1, : nl - 1] = [R.Adown(rad, k + 1, c) + rad.Aup[k, c] for c in range(1, nl)]
        A[k - 1, nl - 1] = rad.Dup[k]
    A[nl - 1, : nl - 1] = rad.Aup[nl, 1:]
    A[nl - 1, nl - 1] = rad.Dup[nl]
    assert np.allclose(A @ f, -1.0, rtol=1e-12)
    assert np.isclose(rbtmoc, ((rad.xlamda - rad.Dmdown) * rbtmat - 1.0) / (rad.xlamda + rad.D0up), rtol=1e-14)


def test_layer_emission_integrals_converge():
    """The trapezoid integrals of the layers' emission (10001 points) agree with a
    finer quadrature of the same integrand to 1e-9."""
    ga = _ga()
    rad = _rad(ga)
    R.radiat(ga, rad)
    # layer 1, upward: int_hm^h1 sigma/2 T(z)^4 exp(-(h1 - z)/zopt) dz / zopt
    z = np.linspace(ga.hm, ga.h[0], 200001)
    T = rad.tabsat[0] - rad.gamma * z
    fine = np.trapezoid(R.SIGOV2 * T ** 4 * np.exp(-(ga.h[0] - z) / rad.zopt[1]), z) / rad.zopt[1]
    zz = np.linspace(ga.hm, ga.h[0], R.NZ)
    TT = rad.tabsat[0] - rad.gamma * zz
    coarse = R._trapin(R.SIGOV2 * TT ** 4 * np.exp(-(ga.h[0] - zz) / rad.zopt[1]), zz[1] - zz[0]) / rad.zopt[1]
    assert abs(coarse - fine) <= 1e-9 * abs(fine)


def test_entrainment_coefficients_and_setup():
    go = Box(385, 385, 3, 10e3, 10e3, 1e-4, 2e-11, [300.0, 1100.0, 2600.0], hm=100.0)
    ga = _ga()
    params, octemp, attemp, rad, xc1, temps = R.coupled_setup(go, ga, yo0=960e3)
    rr = -1.0 / (attemp.rho_cp * attemp.dT_12)
    assert np.isclose(params["bface"], rr * (rad.Aup[0, 0] + R.Adown(rad, 1, 0) - rad.Aup[3, 0]), rtol=1e-14)
    assert np.isclose(params["dface"], rr * (rad.Dup[0] - rad.Dup[3]), rtol=1e-14)
    assert len(params["aface"]) == 2
    # the equilibrium atmospheric temperature is odd about the channel's centre
    assert xc1.shape == (48, 96)
    assert np.allclose(xc1[:, 0], -xc1[::-1, 0], rtol=1e-12)
    assert octemp.T1_rel == 278.0 - temps[0] and octemp.dT_12 == 10.0
