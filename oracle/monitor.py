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


def dP2dx2_bc(p, dx, alpha, cyclic):
    """numerics.F90:427-448."""
    d = np.zeros_like(p)
    fac = 1.0 / (dx * dx)
    d[:, 1:-1] = fac * (p[:, :-2] + p[:, 2:] - 2.0 * p[:, 1:-1])
    if not cyclic:
        d[:, 0] = fac * alpha / (0.5 * alpha + 1.0) * (p[:, 1] - p[:, 0])
        d[:, -1] = -fac * alpha / (0.5 * alpha + 1.0) * (p[:, -1] - p[:, -2])
    else:
        d[:, 0] = fac * (p[:, -2] + p[:, 1] - 2.0 * p[:, 0])
        d[:, -1] = d[:, 0]
    d[0, :] = 0.0
    d[-1, :] = 0.0
    return d


def dP2dy2_bc(p, dy, alpha, cyclic):
    """numerics.F90:450-468."""
    d = np.zeros_like(p)
    fac = 1.0 / (dy * dy)
    d[1:-1, :] = fac * (p[:-2, :] + p[2:, :] - 2.0 * p[1:-1, :])
    d[0, :] = fac * alpha / (0.5 * alpha + 1.0) * (p[1, :] - p[0, :])
    d[-1, :] = -fac * alpha / (0.5 * alpha + 1.0) * (p[-1, :] - p[-2, :])
    if not cyclic:
        d[:, 0] = 0.0
        d[:, -1] = 0.0
    return d


def del2_P_bc(p, dx, dy, alpha, cyclic):
    """del2_P_bc of one layer (numerics.F90:410-425)."""
    return dP2dx2_bc(p, dx, alpha, cyclic) + dP2dy2_bc(p, dy, alpha, cyclic)


def int_P_dx(row, dx):
    """int_P_dx (numerics.F90:206-212)."""
    return (0.5 * row[0] + row[1:-1].sum() + 0.5 * row[-1]) * dx


def energetics(p, pm, wekp, ent1, taux, tauy, dx, dy, fnot, bcco, cyclic, h, gp, ah2, ah4, delek, kt,
               dz_sign, rho):
    """The parts of qg_diagno (qg_monitor.F90:219-383) beyond the state
    diagnostics: Ekman and entrainment averages (:255-264), pken (:273-277),
    utau (:304-306), ah2d, ah4d (:313-327), the jet position and value
    (:334-343), btdg (:352-355), with the lagged velocities from pm.  vg is
    taken (nxp, nyp) like every field it meets (the reference declares it
    (nxt, nyp, nl), :228).  p, pm: (nl, nyp, nxp); wekp, ent1, taux, tauy:
    (nyp, nxp); kt: 0-based k_topo."""
    nl, nyp, nxp = p.shape
    xl, yl = (nxp - 1) * dx, (nyp - 1) * dy
    out = {"ah2d": np.zeros(nl), "ah4d": np.zeros(nl), "val": np.zeros(nl), "pos": np.zeros(nl, dtype=int)}
    out["wepm"] = avg_P(wekp, dx, dy)
    out["wapm"] = avg_P(np.abs(wekp), dx, dy)
    out["entm"] = avg_P(ent1, dx, dy)
    out["enam"] = avg_P(np.abs(ent1), dx, dy)
    if nl > 1:
        eta = dz_sign * (p[0] - p[1]) / gp[0]
        out["pken"] = rho * gp[0] * avg_P(eta * ent1, dx, dy)
    else:
        out["pken"] = 0.0
    ug = np.stack([-dPdy_2bc(pm[k], dy, bcco, cyclic) / fnot for k in range(nl)])
    vg = np.stack([dPdx_2bc(pm[k], dx, bcco, cyclic) / fnot for k in range(nl)])
    ugeos = np.stack([-dPdy_2bc(p[k], dy, bcco, cyclic) / fnot for k in range(nl)])
    vgeos = np.stack([dPdx_2bc(p[k], dx, bcco, cyclic) / fnot for k in range(nl)])
    utaux = int_P_dA(taux * ugeos[0], dx, dy)
    vtauy = int_P_dA(tauy * vgeos[0], dx, dy)
    out["utau"] = rho * (vtauy + utaux)
    for k in range(nl):
        d2u = del2_P_bc(ug[k], dx, dy, bcco, cyclic)
        d4u = del2_P_bc(d2u, dx, dy, bcco, cyclic)
        d2v = del2_P_bc(vg[k], dx, dy, bcco, cyclic)
        d4v = del2_P_bc(d2v, dx, dy, bcco, cyclic)
        u2diss = int_P_dA(ugeos[k] * d2u, dx, dy)
        v2diss = int_P_dA(vgeos[k] * d2v, dx, dy)
        out["ah2d"][k] = -rho * ah2[k] * h[k] * (u2diss + v2diss)
        u4diss = int_P_dA(ugeos[k] * d4u, dx, dy)
        v4diss = int_P_dA(vgeos[k] * d4v, dx, dy)
        out["ah4d"][k] = rho * ah4[k] * h[k] * (u4diss + v4diss)
        ujeto = np.array([abs(int_P_dx(ugeos[k, j], dx)) / xl for j in range(nyp - 1)])   # rows 1..nyt
        out["pos"][k] = int(np.argmax(ujeto)) + 1
        out["val"][k] = float(ujeto.max())
    u2 = int_P_dA(ug[kt] ** 2, dx, dy)
    v2 = int_P_dA(vg[kt] ** 2, dx, dy)
    out["btdg"] = 0.5 * rho * delek * abs(fnot) * (u2 + v2)
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
