"""Pins for the oracle's ocean mixed layer (step_oml, omlsubs.F90:69-119; adv_diff,
mixed_layer.F90:16-95; del2_T, map_T_to_P, c_grid_advection, numerics.F90; the Ekman
velocities of ekman_from_tau, ekman.F90:73-80).

Checked against closed forms on fields where the grid sums are exact, a discrete
eigenfunction, conservation identities that hold for any input, and the Ekman pumping
path already pinned in test_oracle.py."""
import numpy as np
import pytest

from oracle import cpuref as cr
from openqg_amd.forcing import double_gyre_tau

F0 = 1.0e-4


def _cfg(nxp=17, nyp=13, cyclic=False, dx=10e3, tau_sign=1):
    return dict(nxp=nxp, nyp=nyp, nl=3, cyclic=cyclic, dx=dx, dy=dx, fnot=F0, beta=2e-11, h=[300.0, 1100.0, 2600.0],
                gp=[0.05, 0.025], ah2=[0.0] * 3, ah4=[1e10] * 3, bcco=0.5, delek=2.0, dz_sign=-1, tau_sign=tau_sign)


def _oml(**kw):
    o = dict(hm=100.0, T1_rel=-2.0, dT_12=10.0, rho_cp=4.0e6, d2=380.0, d4=4.0e10, ycexp=1.0, nval=-3.0, sval=3.0,
             fluxn=False, fluxs=True)
    o.update(kw)
    return o


def _trap(f, dx):
    w = np.ones(f.shape)
    w[0, :] *= 0.5
    w[-1, :] *= 0.5
    w[:, 0] *= 0.5
    w[:, -1] *= 0.5
    return float((w * f).sum()) * dx * dx


def test_del2_T_closed_forms():
    """Interior: Del^2 of a quadratic is 2(a+b)/dx^2.  No-flux walls take a zero outer
    difference; flux walls the difference to the prescribed value; in a channel a discrete
    Fourier mode in x is an eigenfunction, 2(cos(2 pi k/nxt) - 1)/dx^2."""
    cfg = _cfg()
    nxt, nyt, dx = 16, 12, 10e3
    I = np.arange(1, nxt + 1)[None, :]
    J = np.arange(1, nyt + 1)[:, None]
    a, b = 0.3, -1.1
    t = a * I ** 2 + b * J ** 2 + 0.0 * I
    d = cr.del2_T(cfg, t)
    assert np.allclose(d[1:-1, 1:-1], 2 * (a + b) / dx ** 2, rtol=1e-12, atol=0)
    c = 0.7
    lin = c * I + 0.0 * J
    d = cr.del2_T(cfg, lin)
    assert np.allclose(d[:, 1:-1], 0.0, atol=1e-12 * c / dx ** 2)
    assert np.allclose(d[:, 0], c / dx ** 2, rtol=1e-13) and np.allclose(d[:, -1], -c / dx ** 2, rtol=1e-13)
    # flux walls in y: dt_y(1) = t(1) - sval, dt_y(nyt+1) = nval - t(nyt)
    u = np.full((nyt, nxt), 2.0)
    d = cr.del2_T(cfg, u, north_flux=True, nval=5.0, south_flux=True, sval=-1.0)
    assert np.allclose(d[0], (0.0 - (2.0 + 1.0)) / dx ** 2, rtol=1e-14)
    assert np.allclose(d[-1], (5.0 - 2.0) / dx ** 2, rtol=1e-14)
    assert np.all(d[1:-1] == 0.0)
    cc = _cfg(cyclic=True)
    for k in (1, 3):
        m = np.cos(2 * np.pi * k * (I - 1) / nxt) + 0.0 * J
        d = cr.del2_T(cc, m)
        lam = 2 * (np.cos(2 * np.pi * k / nxt) - 1) / dx ** 2
        assert np.abs(d[1:-1] - lam * m[1:-1]).max() <= 1e-12 * abs(lam)


def test_ekman_velocities_closed_forms_and_pumping():
    """Uniform taux: vek = -tau_sign taux/(f0 hm), no pumping.  taux linear in y: wekt =
    -slope/f0.  In general map_T_to_P(wekt) is the Ekman pumping wekp of the pinned
    ekman_from_tau + map_T_to_P path."""
    cfg = _cfg(tau_sign=-1)
    nyp, nxp, dx, hm = 13, 17, 10e3, 80.0
    tx = np.full((nyp, nxp), 2e-4)
    ty = np.zeros((nyp, nxp))
    u, v, w = cr.ekman_uvw(cfg, hm, tx, ty)
    assert np.allclose(v, 2e-4 / (F0 * hm), rtol=1e-14) and np.all(u == 0) and np.all(w == 0)
    y = np.arange(nyp)[:, None] * dx
    s = 3e-11
    u, v, w = cr.ekman_uvw(cfg, hm, s * y + 0 * tx, ty)
    assert np.allclose(w, -s / F0, rtol=1e-10)
    tx, ty = double_gyre_tau(nyp, nxp, dx, 0.1)
    ty = 0.2 * np.roll(tx, 2, axis=0) + 0.05 * np.sin(2 * np.pi * np.arange(nxp) / (nxp - 1))[None, :] * tx.max()
    _, _, w = cr.ekman_uvw(cfg, hm, tx, ty)
    wekp, _, _ = cr.ekman_wekp(cfg, tx, ty)
    assert np.abs(cr.map_T_to_P(cfg, w) - wekp).max() <= 1e-14 * np.abs(wekp).max()


def test_advection_uniform_flow_and_conservation():
    """p linear in y gives a uniform zonal geostrophic flow u = -(ycexp/f0) dp/dy; for a
    temperature linear in x the C-grid flux divergence is -u dT/dx away from the walls.  For
    any fields, with no-flux walls (no normal velocity), the fluxes telescope: the area sum
    of the advective tendency is zero."""
    cfg = _cfg()
    nyp, nxp, dx = 13, 17, 10e3
    nyt, nxt = nyp - 1, nxp - 1
    o = _oml(d2=0.0, d4=0.0, fluxs=False, fluxn=False)
    g = 2.0e-3
    p = g * (np.arange(nyp)[:, None] * dx) + 0 * np.zeros((nyp, nxp))
    c = 1e-5
    T = c * (np.arange(nxt)[None, :] * dx) + np.zeros((nyt, nxt))
    z = np.zeros
    ad = cr.adv_diff(cfg, o, p, z((nyt, nxp)), z((nyp, nxt)), T, T)
    u = -(1.0 / F0) * g
    assert np.allclose(ad[:, 1:-1], -u * c, rtol=1e-10)
    # p linear in x: uniform meridional flow v = (ycexp/f0) dp/dx; T linear in y
    p = g * (np.arange(nxp)[None, :] * dx) + np.zeros((nyp, nxp))
    T = c * (np.arange(nyt)[:, None] * dx) + np.zeros((nyt, nxt))
    ad = cr.adv_diff(cfg, o, p, z((nyt, nxp)), z((nyp, nxt)), T, T)
    v = (1.0 / F0) * g
    assert np.allclose(ad[1:-1, :], -v * c, rtol=1e-10)
    rng = np.random.default_rng(3)
    p = rng.standard_normal((nyp, nxp))
    uek, vek = 1e-2 * rng.standard_normal((nyt, nxp)), 1e-2 * rng.standard_normal((nyp, nxt))
    T = rng.standard_normal((nyt, nxt))
    ad = cr.adv_diff(cfg, o, p, uek, vek, T, rng.standard_normal((nyt, nxt)))
    scale = np.abs(ad).max() * ad.size
    assert abs(ad.sum()) <= 1e-13 * scale
    # the Del^4 term alone also conserves (zero-gradient walls); Del^2 with flux walls does not
    o2 = _oml(d2=0.0, d4=1.0, fluxs=False, fluxn=False, ycexp=0.0)
    ad4 = cr.adv_diff(cfg, o2, 0 * p, 0 * uek, 0 * vek, 0 * T, T)
    assert abs(ad4.sum()) <= 1e-13 * np.abs(ad4).max() * ad4.size


def test_oml_heat_budget_and_convection():
    """No flow, no diffusion, a uniform layer: sst_new = sstm + tdto F/(rho_cp hm); when that
    falls below T1_rel the convective correction restores T1_rel (omlsubs.F90:98-114), and the
    entrainment, uniform before its mean is removed, is zero."""
    cfg = _cfg()
    nyp, nxp = 13, 17
    nyt, nxt = nyp - 1, nxp - 1
    z = np.zeros
    for F, T0 in ((150.0, 1.0), (-2.0e4, -1.95)):
        o = _oml(d2=0.0, d4=0.0)
        tdto = 2400.0
        sst, sstm, ent = cr.oml_step(cfg, o, z((nyp, nxp)), z((nyp, nxp)), z((nyp, nxp)), np.full((nyt, nxt), F), tdto,
                                     np.full((nyt, nxt), T0 + 0.5), np.full((nyt, nxt), T0))
        new = T0 + tdto * (F / (o["rho_cp"] * o["hm"]))
        exp = max(new, o["T1_rel"])
        assert np.allclose(sst, exp, rtol=1e-14) and np.all(sstm == T0 + 0.5)
        assert np.abs(ent).max() <= 1e-12 * (o["hm"] / (tdto * o["dT_12"])) * abs(o["T1_rel"] - new)


def test_oml_entrainment_has_zero_net_mass():
    """The entrainment is mapped from T points after its mean is removed, and the trapezoid
    integral of map_T_to_P(t) is dx dy times the sum of t, so the net entrainment vanishes
    (no net heat flux into the deep ocean, omlsubs.F90:105-110)."""
    for cyc in (False, True):
        cfg = _cfg(cyclic=cyc)
        nyp, nxp, dx = 13, 17, 10e3
        nyt, nxt = nyp - 1, nxp - 1
        rng = np.random.default_rng(5 + cyc)
        t = rng.standard_normal((nyt, nxt))
        m = cr.map_T_to_P(cfg, t)
        if not cyc:
            assert abs(_trap(m, dx) - t.sum() * dx * dx) <= 1e-13 * np.abs(t).sum() * dx * dx
        tx, ty = 1e-4 * rng.standard_normal((nyp, nxp)), 1e-4 * rng.standard_normal((nyp, nxp))
        p1 = 10.0 * rng.standard_normal((nyp, nxp))
        if cyc:
            for a in (tx, ty, p1):
                a[:, -1] = a[:, 0]
        sst = rng.standard_normal((nyt, nxt))
        _, _, ent = cr.oml_step(cfg, _oml(), p1, tx, ty, 50 * rng.standard_normal((nyt, nxt)), 2400.0, sst,
                                sst + 0.1 * rng.standard_normal((nyt, nxt)))
        if not cyc:
            assert abs(_trap(ent, dx)) <= 1e-13 * np.abs(ent).sum() * dx * dx
        assert np.abs(ent).max() > 0
