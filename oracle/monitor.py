"""Test infrastructure only (part of the oracle): a numpy restatement of the
diagnostics of the reference's qg_monitor / valids that the device
diagnostics reproduce, used by tests/ to compare against, never by the
openqg_amd package.  Arrays are (nyp, nxp) per layer, the transpose of the
reference's Fortran (nxp, nyp).
"""
from __future__ import annotations

import numpy as np


def int_P_dA(f, dx, dy):
    """int_P_dA_2d (glam/numerics.F90:188-204): all points, minus half the
    N/S and E/W rows, plus a quarter of the corners, times dx dy."""
    s = f.sum()
    s = s - 0.5 * (f[0, :].sum() + f[-1, :].sum())
    s = s - 0.5 * (f[:, 0].sum() + f[:, -1].sum())
    s = s + 0.25 * (f[0, 0] + f[-1, 0] + f[0, -1] + f[-1, -1])
    return s * dx * dy


def avg_P(f, dx, dy):
    """avg_P (numerics.F90:214-221)."""
    nyp, nxp = f.shape
    return int_P_dA(f, dx, dy) / (((nxp - 1) * dx) * ((nyp - 1) * dy))


def dPdx_2bc(p, dx, alpha, cyclic):
    """numerics.F90:470-492."""
    d = np.zeros_like(p)
    d[:, 1:-1] = (p[:, 2:] - p[:, :-2]) / (2.0 * dx)
    if not cyclic:
        d[:, 0] = (p[:, 1] - p[:, 0]) / (dx * (0.5 * alpha + 1.0))
        d[:, -1] = (p[:, -1] - p[:, -2]) / (dx * (0.5 * alpha + 1.0))
    else:
        d[:, 0] = (p[:, 1] - p[:, -2]) / (2.0 * dx)
        d[:, -1] = d[:, 0]
    d[0, :] = 0.0
    d[-1, :] = 0.0
    return d


def dPdy_2bc(p, dy, alpha, cyclic):
    """numerics.F90:494-513."""
    d = np.zeros_like(p)
    d[1:-1, :] = (p[2:, :] - p[:-2, :]) / (2.0 * dy)
    d[0, :] = (p[1, :] - p[0, :]) / (dy * (0.5 * alpha + 1.0))
    d[-1, :] = (p[-1, :] - p[-2, :]) / (dy * (0.5 * alpha + 1.0))
    if not cyclic:
        d[:, 0] = 0.0
        d[:, -1] = 0.0
    return d


def diagnose(q, p, dx, dy, fnot, bcco, cyclic, h, gp, dz_sign, rho, dt):
    """The state-only parts of qg_diagno (qg_monitor.F90:219-383), scan_cfl's
    maxima (:507-569, with dPdx/dPdy of numerics.F90:291-309) and the count of
    non-finite p and q.  q, p: (nl, nyp, nxp)."""
    nl, nyp, nxp = p.shape
    out = {k: np.zeros(nl) for k in ("ekei", "pavg", "qavg", "pmin", "pmax", "qmin", "qmax", "sfmin", "sfmax",
                                      "circ", "umax", "vmax", "cflu", "cflv")}
    out["etam"] = np.zeros(nl - 1)
    out["et2m"] = np.zeros(nl - 1)
    for k in range(nl):
        u = -dPdy_2bc(p[k], dy, bcco, cyclic) / fnot
        v = dPdx_2bc(p[k], dx, bcco, cyclic) / fnot
        uke = int_P_dA(u * u, dx, dy)
        vke = int_P_dA(v * v, dx, dy)
        out["ekei"][k] = 0.5 * rho * h[k] * (uke + vke)
        out["pavg"][k] = avg_P(p[k], dx, dy)
        out["qavg"][k] = avg_P(q[k], dx, dy)
        out["pmin"][k], out["pmax"][k] = np.nanmin(p[k]), np.nanmax(p[k])
        out["qmin"][k], out["qmax"][k] = np.nanmin(q[k]), np.nanmax(q[k])
        out["circ"][k] = 1.0e-6 * ((p[k, 0, 0] - p[k, -1, 0]) / fnot)
        poref = p[k, 0, 0] if fnot > 0 else p[k, -1, 0]
        lo = min(out["pmin"][k] / fnot, out["pmax"][k] / fnot)
        hi = max(out["pmin"][k] / fnot, out["pmax"][k] / fnot)
        out["sfmin"][k] = 1.0e-6 * (h[k] * (lo - poref / fnot))
        out["sfmax"][k] = 1.0e-6 * (h[k] * (hi - poref / fnot))
        ug = -((1.0 / dy) * (p[k, 1:, :] - p[k, :-1, :])) / fnot
        vg = ((1.0 / dx) * (p[k, :, 1:] - p[k, :, :-1])) / fnot
        out["umax"][k] = np.nanmax(np.abs(ug))
        out["vmax"][k] = np.nanmax(np.abs(vg))
        out["cflu"][k] = out["umax"][k] * dt / dx
        out["cflv"][k] = out["vmax"][k] * dt / dx
    for k in range(nl - 1):
        eta = dz_sign * (p[k] - p[k + 1]) / gp[k]
        out["etam"][k] = avg_P(eta, dx, dy)
        out["et2m"][k] = avg_P(eta * eta, dx, dy)
    out["ctot"] = float(out["circ"].sum())
    out["nonfinite"] = float(np.count_nonzero(~np.isfinite(p)) + np.count_nonzero(~np.isfinite(q)))
    return out


def check_continuity(int_p, dpi, gp, dz_sign, xl, yl, tdt):
    """check_continuity (qg_monitor.F90:571-604), called for cyclic basins
    (:378-381): per interface k, ermas = dz_sign (int_p_k - int_p_k+1) - dpi_k
    and emfr = 2 ermas / (|est1| + |est2|) when that sum exceeds
    1e-13 xl yl tdt gp_k, else 0.  Returns (ermas, emfr)."""
    nl = len(int_p)
    ermas, emfr = np.zeros(nl - 1), np.zeros(nl - 1)
    for k in range(nl - 1):
        est1 = dz_sign * (int_p[k] - int_p[k + 1])
        est2 = dpi[k]
        edif = est1 - est2
        esum = abs(est1) + abs(est2)
        ermas[k] = edif
        emfr[k] = 2.0 * edif / esum if esum > 1.0e-13 * xl * yl * tdt * gp[k] else 0.0
    return ermas, emfr
