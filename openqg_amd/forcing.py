"""Wind-stress forcing producer: Ekman pumping at p points.

ekman_from_tau (src/coupler/ekman.F90:53-83) with map_P_to_x_face /
map_P_to_y_face (numerics.F90:271-289) and map_T_to_P (numerics.F90:64-107).
For ocean-only runs the stress is static, so wekp is made once at set-up;
this is input preparation, not part of the per-step path.
"""
from __future__ import annotations

import numpy as np


def _int_P_dx(row: np.ndarray, dx: float) -> float:
    return (0.5 * row[0] + row[1:-1].sum() + 0.5 * row[-1]) * dx


def ekman_wekp(nxp: int, nyp: int, dx: float, dy: float, fnot: float, cyclic: bool,
               taux: np.ndarray, tauy: np.ndarray):
    """Return (wekp (nyp,nxp), txis, txin) from kinematic stress (m^2/s^2) at p points."""
    taux = np.asarray(taux, dtype=np.float64)
    tauy = np.asarray(tauy, dtype=np.float64)
    txis = txin = 0.0
    if cyclic:
        txis = 0.5 * _int_P_dx(taux[0] + taux[1], dx)
        txin = 0.5 * _int_P_dx(taux[-2] + taux[-1], dx)
    taux_t = 0.5 * (taux[:, 1:] + taux[:, :-1])     # (nyp, nxp-1)  x faces
    tauy_t = 0.5 * (tauy[1:, :] + tauy[:-1, :])     # (nyp-1, nxp)  y faces
    wekt = ((1.0 / (dx * fnot)) * (tauy_t[:, 1:] - tauy_t[:, :-1])
            - (1.0 / (dy * fnot)) * (taux_t[1:, :] - taux_t[:-1, :]))   # (nyt, nxt)
    t = wekt
    w = np.zeros((nyp, nxp))
    w[1:-1, 1:-1] = 0.25 * (t[1:, 1:] + t[1:, :-1] + t[:-1, 1:] + t[:-1, :-1])
    w[0, 1:-1] = 0.5 * (t[0, :-1] + t[0, 1:])
    w[-1, 1:-1] = 0.5 * (t[-1, :-1] + t[-1, 1:])
    if cyclic:
        w[1:-1, 0] = 0.25 * (t[1:, 0] + t[1:, -1] + t[:-1, 0] + t[:-1, -1])
        w[1:-1, -1] = w[1:-1, 0]
        w[0, 0] = 0.5 * (t[0, 0] + t[0, -1])
        w[-1, 0] = 0.5 * (t[-1, 0] + t[-1, -1])
        w[0, -1] = w[0, 0]
        w[-1, -1] = w[-1, 0]
    else:
        w[1:-1, 0] = 0.5 * (t[:-1, 0] + t[1:, 0])
        w[1:-1, -1] = 0.5 * (t[:-1, -1] + t[1:, -1])
        w[0, 0] = t[0, 0]
        w[0, -1] = t[0, -1]
        w[-1, 0] = t[-1, 0]
        w[-1, -1] = t[-1, -1]
    return w, txis, txin


def ekman_wekp_adjoint(nxp: int, nyp: int, dx: float, dy: float, fnot: float, lam_w: np.ndarray):
    """Transpose of (taux, tauy) -> wekp of ekman_wekp for a box: given the
    adjoint of wekp (nyp, nxp), the adjoints of the kinematic stresses
    (taux, tauy), each (nyp, nxp).  Every line reverses one line above."""
    lw = np.asarray(lam_w, dtype=np.float64)
    lt = np.zeros((nyp - 1, nxp - 1))
    # map_T_to_P (box)
    q = 0.25 * lw[1:-1, 1:-1]
    lt[1:, 1:] += q
    lt[1:, :-1] += q
    lt[:-1, 1:] += q
    lt[:-1, :-1] += q
    lt[0, :-1] += 0.5 * lw[0, 1:-1]
    lt[0, 1:] += 0.5 * lw[0, 1:-1]
    lt[-1, :-1] += 0.5 * lw[-1, 1:-1]
    lt[-1, 1:] += 0.5 * lw[-1, 1:-1]
    lt[:-1, 0] += 0.5 * lw[1:-1, 0]
    lt[1:, 0] += 0.5 * lw[1:-1, 0]
    lt[:-1, -1] += 0.5 * lw[1:-1, -1]
    lt[1:, -1] += 0.5 * lw[1:-1, -1]
    lt[0, 0] += lw[0, 0]
    lt[0, -1] += lw[0, -1]
    lt[-1, 0] += lw[-1, 0]
    lt[-1, -1] += lw[-1, -1]
    # wekt = (tauy_t[:,1:] - tauy_t[:,:-1])/(dx f0) - (taux_t[1:,:] - taux_t[:-1,:])/(dy f0)
    ltauy_t = np.zeros((nyp - 1, nxp))
    ltaux_t = np.zeros((nyp, nxp - 1))
    ltauy_t[:, 1:] += lt / (dx * fnot)
    ltauy_t[:, :-1] -= lt / (dx * fnot)
    ltaux_t[1:, :] -= lt / (dy * fnot)
    ltaux_t[:-1, :] += lt / (dy * fnot)
    # face averages
    ltaux = np.zeros((nyp, nxp))
    ltauy = np.zeros((nyp, nxp))
    ltaux[:, 1:] += 0.5 * ltaux_t
    ltaux[:, :-1] += 0.5 * ltaux_t
    ltauy[1:, :] += 0.5 * ltauy_t
    ltauy[:-1, :] += 0.5 * ltauy_t
    return ltaux, ltauy


def double_gyre_tau(nyp: int, nxp: int, dy: float, tau0: float, rho0: float = 1000.0):
    """tau_x = tau0 (1 - cos(2 pi y / L)) / rho0 with y = (j-1) dy, L = (nyp-1) dy; tau_y = 0."""
    y = np.arange(nyp) * dy
    L = (nyp - 1) * dy
    tx = tau0 * (1.0 - np.cos(2.0 * np.pi * y / L)) / rho0
    return np.repeat(tx[:, None], nxp, axis=1), np.zeros((nyp, nxp))
