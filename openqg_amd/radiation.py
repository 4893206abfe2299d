"""The radiation balance of OpenQG's coupled model, computed at set-up on the host.

`load_rad` and `radiat` restate src/mixed/radsubs.F90:63-95 and :116-676: the
mean radiative state (the atmospheric and ocean mixed-layer mean temperatures
that close the balance), the linearised perturbation coefficients Aup, Adown,
Dup, D0up, Dmdown the coupled step reads, and the radiative-balance multipliers
rbetat, rbtmat, rbtmoc.  `entrain_coeff` is init_entrain_coeff
(src/mixed/amlsubs.F90:93-129), `xc1ast` is set_xc1ast's first option
(amlsubs.F90:131-182) and `coupler_params` gathers what the device coupler
takes (openqg_amd.coupled.Coupler).  Set-up code: none of it runs per step.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mixed import Temp, init_temp_from_rad, zonal_temp
from .qg import Box

STEFAN = 5.67040e-8          # constants.F90:20
SIGOV2 = 0.5 * STEFAN
NZ = 10001                   # quadrature points of the layer integrals (radsubs.F90:208)
TMBTOL, NITMAX = 1.0e-13, 200


@dataclass
class Radiation:
    """radiate_type (radsubs.F90:14-53); Aup[k][l] k = 0..nl, Adown[k][l] k = 1..nl at row k-1, l = 0..nl-1."""

    fsbar: float
    fspamp: float
    zopt: np.ndarray            # (nl+1,) optical depths, mixed layer first
    gamma: float
    xlamda: float
    tabsat: np.ndarray          # (nl,) absolute layer temperatures
    fspco: float = 0.0
    Aup: np.ndarray = field(default=None)
    Adown: np.ndarray = field(default=None)
    Dup: np.ndarray = field(default=None)
    D0up: float = 0.0
    Dmdown: float = 0.0
    rbetat: np.ndarray = field(default=None)


def load_rad(ga: Box, fsbar, fspamp, zopt, gamma, xlamda, T_abs) -> Radiation:
    """load_rad (radsubs.F90:63-95): the signed perturbation amplitude takes f0's sign."""
    if fspamp < 0.0:
        raise ValueError("fspamp must be >= 0 (radsubs.F90:88-92)")
    return Radiation(float(fsbar), float(fspamp), np.asarray(zopt, dtype=float), float(gamma), float(xlamda),
                     np.asarray(T_abs, dtype=float), fspco=math.copysign(float(fspamp), ga.fnot))


def _trans(thickness, opt_depth):
    return np.exp(-thickness / opt_depth)


def _trapin(f, delz):
    """trapin (radsubs.F90:664-672): the trapezoid rule on NZ equal steps."""
    return delz * (0.5 * f[0] + f[1:-1].sum() + 0.5 * f[-1])


def _B(T0, z, gamma):
    return SIGOV2 * (T0 - z * gamma) ** 4


def radiat(ga: Box, rad: Radiation):
    """radiat (radsubs.F90:116-153).  Fills rad's coefficients; returns
    (rbtmat, rbtmoc, ocn_tmbar, atm_tmbar)."""
    nl, hm, h = ga.nl, ga.hm, np.asarray(ga.h, dtype=float)
    zopt, gamma = rad.zopt, rad.gamma
    # compute_layer_trans (:155-170)
    tauk = np.empty(nl + 1)
    tauk[0] = _trans(hm, zopt[0])
    tauk[1] = _trans(h[0] - hm, zopt[1])
    for k in range(2, nl + 1):
        tauk[k] = _trans(h[k - 1], zopt[k])
    # compute_mean_radiation (:179-246): up and down emission of each layer
    uprad, dnrad = np.empty(nl + 1), np.empty(nl + 1)   # 1-based
    i = np.arange(NZ, dtype=float)
    hbot, htop = hm, h[0]
    for k in range(1, nl + 1):
        if k == 1:
            delz = (htop - hbot) / (NZ - 1)
        else:
            hbot, htop = htop, htop + h[k - 1]
            delz = h[k - 1] / (NZ - 1)
        zz = hbot + i * delz
        T = rad.tabsat[k - 1] - gamma * zz
        fup = SIGOV2 * T ** 4 * _trans(htop - zz, zopt[k])
        fdn = SIGOV2 * T ** 4 * _trans(zz - hbot, zopt[k])
        uprad[k] = _trapin(fup, delz) / zopt[k]
        dnrad[k] = _trapin(fdn, delz) / zopt[k]
    # compute_ml_mean_temp (:248-303): the atmospheric mixed layer's mean temperature
    rhstat = 0.0
    for k in range(1, nl + 1):
        rhstat = rhstat * tauk[k] + uprad[k]
    rhstat = (-rhstat - rad.fsbar) / np.prod(tauk[1:])
    atm_tmbar = 300.0
    delz = hm / (NZ - 1)
    zz = i * delz
    for it in range(NITMAX + 1):
        fup = (atm_tmbar - gamma * zz) ** 4 * _trans(hm - zz, zopt[0])
        uprad_m = _trapin(fup, delz) * SIGOV2 / zopt[0]
        deltm = 0.25 * (rhstat - uprad_m) * atm_tmbar / uprad_m
        atm_tmbar = atm_tmbar + 0.75 * deltm
        if abs(deltm) <= TMBTOL:
            break
    else:
        raise RuntimeError("atmospheric mixed-layer temperature did not converge (radsubs.F90:291-296)")
    # compute_oml_mean_temp (:305-336)
    rhstoc = rad.xlamda * atm_tmbar + _B(atm_tmbar, 0.0, 0.0) - rad.fsbar
    ocn_tmbar = atm_tmbar
    for it in range(NITMAX + 1):
        old = ocn_tmbar
        ocn_tmbar = rhstoc / (rad.xlamda + STEFAN * old ** 3)
        if abs(ocn_tmbar - old) <= TMBTOL:
            break
    else:
        raise RuntimeError("ocean mixed-layer temperature did not converge (radsubs.F90:325-331)")
    # compute_mean_flux (:346-373)
    Fup = np.empty(nl + 1)
    Fup[0] = uprad_m
    for k in range(1, nl + 1):
        Fup[k] = Fup[k - 1] * tauk[k] + uprad[k]
    Fdn = np.empty(nl + 1)   # 1-based
    Fdn[nl] = -dnrad[nl]
    for k in range(nl - 1, 0, -1):
        Fdn[k] = Fdn[k + 1] * tauk[k] - dnrad[k]
    # compute_perturb_rad_params (:430-521)
    Z = np.cumsum(h)
    Btop = np.empty(nl + 1)
    Bbot = np.empty(nl + 1)
    Btop[0] = _B(atm_tmbar, hm, gamma)
    for k in range(1, nl + 1):
        Btop[k] = _B(rad.tabsat[k - 1], Z[k - 1], gamma)
    Bbot[0] = _B(atm_tmbar, 0.0, gamma)
    Bbot[1] = _B(rad.tabsat[0], hm, gamma)
    for k in range(2, nl + 1):
        Bbot[k] = _B(rad.tabsat[k - 1], Z[k - 2], gamma)
    Aup = np.zeros((nl + 1, nl))
    Adn = np.zeros((nl + 1, nl))   # row k = 1..nl used
    Aup[0, 0] = (Btop[0] - Fup[0]) / zopt[0]
    for k in range(1, nl):
        Aup[k, k] = (-tauk[k] * Fup[k - 1] - uprad[k] + Btop[k]) / zopt[k]
    for k in range(1, nl + 1):
        Aup[k, k - 1] = tauk[k] * (Aup[k - 1, k - 1] + Fup[k - 1] / zopt[k] - Bbot[k] / zopt[k])
    for k in range(2, nl + 1):
        for l in range(0, k - 1):
            Aup[k, l] = Aup[k - 1, l] * tauk[k]
    rad.D0up = 4.0 * STEFAN * ocn_tmbar ** 3
    zz = i * delz
    fup = (atm_tmbar - gamma * zz) ** 3 * np.exp(-(hm - zz) / zopt[0])
    Dup = np.empty(nl + 1)
    Dup[0] = 2.0 * STEFAN * _trapin(fup, delz) / zopt[0]
    for k in range(1, nl + 1):
        Dup[k] = Dup[k - 1] * tauk[k]
    for k in range(1, nl):
        Adn[k, k - 1] = (Fdn[k + 1] * tauk[k] - dnrad[k] + Bbot[k]) / zopt[k]
    Adn[nl, nl - 1] = (Bbot[nl] - dnrad[nl]) / zopt[nl]
    for k in range(nl - 1, 0, -1):
        for l in range(k + 1, nl):
            Adn[k, l] = Adn[k + 1, l] * tauk[k]
        Adn[k, k] = tauk[k] * (Adn[k + 1, k] - Fdn[k + 1] / zopt[k] - Btop[k] / zopt[k])
    rad.Dmdown = -2.0 * STEFAN * atm_tmbar ** 3
    rad.Aup, rad.Adown, rad.Dup = Aup, Adn[1:], Dup
    # compute_rad_bal_coeffs (:613-662): Fs' multipliers of eta and the mixed-layer temperatures
    A = np.zeros((nl, nl))
    for c in range(1, nl):
        A[0, c - 1] = Adn[1, c]
    A[0, nl - 1] = Dup[0]
    for k in range(2, nl):
        for c in range(1, nl):
            A[k - 1, c - 1] = Adn[k + 1, c] + Aup[k, c]
        A[k - 1, nl - 1] = Dup[k]
    for c in range(1, nl):
        A[nl - 1, c - 1] = Aup[nl, c]
    A[nl - 1, nl - 1] = Dup[nl]
    f = np.linalg.solve(A, -np.ones(nl))
    rad.rbetat = f[: nl - 1]
    rbtmat = float(f[nl - 1])
    rbtmoc = ((rad.xlamda - rad.Dmdown) * rbtmat - 1.0) / (rad.xlamda + rad.D0up)
    return rbtmat, rbtmoc, float(ocn_tmbar), float(atm_tmbar)


def Adown(rad: Radiation, k: int, l: int) -> float:
    """Adown(k, l) with the reference's 1-based k."""
    return float(rad.Adown[k - 1, l])


def entrain_coeff(nl: int, attemp: Temp, rad: Radiation):
    """init_entrain_coeff (amlsubs.F90:93-129): (aface(1..nl-1), bface, dface)."""
    rr = -1.0 / (attemp.rho_cp * attemp.dT_12)
    aface = [rr * (0.0 + Adown(rad, 1, l) - rad.Aup[nl, l]) for l in range(1, nl)]
    bface = rr * (rad.Aup[0, 0] + Adown(rad, 1, 0) - rad.Aup[nl, 0])
    dface = rr * (rad.Dup[0] + 0.0 - rad.Dup[nl])
    return aface, float(bface), float(dface)


def xc1ast(ga: Box, attemp: Temp) -> np.ndarray:
    """set_xc1ast's radiative-equilibrium option (amlsubs.F90:145-150): (nyt, nxt)."""
    rows = np.array([zonal_temp(ga, attemp, j) for j in range(1, ga.nyt + 1)])
    return rows[:, None] + np.zeros((1, ga.nxt))


def coupled_setup(go: Box, ga: Box, *, fsbar=-210.0, fspamp=80.0, zopt=(200.0, 3.0e4, 3.0e4, 4.0e4), gamma=1.0e-2,
                  xlamda=35.0, T_abs=(330.0, 340.0, 350.0), T1_ocn=278.0, T2_ocn=268.0, rho_o=1.0e3, cp_o=4.0e3,
                  rho_a=1.0, cp_a=1.0e3, hmamin=100.0, hmadmp=0.15, xcexp=1.0, at2d=5.0e4, at4d=3.0e14,
                  ahmd=2.0e5, cdat=1.3e-3, tau_udiff=False, xo0=0.0, yo0=0.0, xa0=0.0, ya0=0.0):
    """The set-up of init_mixed_and_qg (openqg.F90:190-218) for the double-gyre pair, with
    the defaults of data/basin/atm_dg.cdl and ocn_dg.cdl: returns (params for Coupler,
    ocean Temp, atmosphere Temp, Radiation, xc1ast, (ocn_tmbar, atm_tmbar, rbtmat, rbtmoc))."""
    rad = load_rad(ga, fsbar, fspamp, zopt, gamma, xlamda, T_abs)
    rbtmat, rbtmoc, ocn_tmbar, atm_tmbar = radiat(ga, rad)
    attemp = init_temp_from_rad(ga, ga, rad.tabsat[0], rad.tabsat[1], atm_tmbar, rbtmat, rad.fspco, rho_a, cp_a)
    octemp = init_temp_from_rad(go, ga, T1_ocn, T2_ocn, ocn_tmbar, rbtmoc, rad.fspco, rho_o, cp_o)
    aface, bface, dface = entrain_coeff(ga.nl, attemp, rad)
    params = dict(xo0=xo0, yo0=yo0, xa0=xa0, ya0=ya0, hmo=go.hm, hma=ga.hm, rho_o=rho_o, rho_a=rho_a, cdat=cdat,
                  tau_udiff=tau_udiff, D0up=rad.D0up, Dmdown=rad.Dmdown, Dup0=float(rad.Dup[0]), xlamda=rad.xlamda,
                  fspco=rad.fspco, Aup00=float(rad.Aup[0, 0]), Adown10=Adown(rad, 1, 0), Adown11=Adown(rad, 1, 1),
                  hmamin=hmamin, hmadmp=hmadmp, xcexp=xcexp, T1_rel=attemp.T1_rel, dT_12=attemp.dT_12,
                  rho_cp=attemp.rho_cp, at2d=at2d, at4d=at4d, ahmd=ahmd, aface=aface, bface=bface, dface=dface)
    return params, octemp, attemp, rad, xc1ast(ga, attemp), (ocn_tmbar, atm_tmbar, rbtmat, rbtmoc)
