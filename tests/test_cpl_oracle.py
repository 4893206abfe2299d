"""Pins for the oracle's ocean-atmosphere coupling: the coupling grid (grid.F90:37-58),
the second-order derivative stencils with boundary conditions (numerics.F90:427-513),
bicubic interpolation (bicubic.F90), the coupled wind stress (windstress.F90:102-316),
bilinear interpolation (numerics.F90:528-589), the heat-flux forcing
(forcing.F90:18-99) and the atmospheric mixed layer (amlsubs.F90:184-326).

Each is checked against what the mathematics fixes: exactness of the stencils on
polynomials and Fourier modes, exact reproduction of bicubic polynomials by the bicubic
interpolant (which fixes the Hermite matrix, the corner-data order and the cell
placement), the quadratic drag law that the stress formula solves (|u|^2 =
(|tau|/cd)(1 + k^2 |tau|/cd) and the Ekman turning), exact bilinear interpolation of
linear fields, closed forms on uniform states, and the conservation of the heat flux
passed from the ocean grid to the atmosphere's."""
import numpy as np
import pytest

from oracle import cpuref as cr

F0 = 1.0e-4


def _ocean(n=33, dx=10e3):
    return dict(nxp=n, nyp=n, nl=3, cyclic=False, dx=dx, dy=dx, fnot=F0, beta=2e-11, h=[300.0, 1100.0, 2600.0],
                gp=[0.05, 0.025], ah2=[0.0] * 3, ah4=[1e10] * 3, bcco=0.5, delek=2.0, dz_sign=-1, tau_sign=1)


def _atmos(nxp=9, nyp=9, dx=80e3):
    return dict(nxp=nxp, nyp=nyp, nl=3, cyclic=True, dx=dx, dy=dx, fnot=F0, beta=2e-11, h=[2000.0, 3000.0, 4000.0],
                gp=[1.2, 0.4], ah2=[0.0] * 3, ah4=[2e14] * 3, bcco=1.0, delek=0.0, dz_sign=1, tau_sign=-1)


def _grid(tau_udiff=False, **kw):
    co, ca = _ocean(**kw), _atmos()
    g = cr.cpl(co, ca, hmo=100.0, hma=1000.0, rho_o=1000.0, rho_a=1.0, cdat=1.3e-3, tau_udiff=tau_udiff)
    g.yo0 = (g.ny1 - 1) * ca["dx"]   # the ocean box starts at atmos p row ny1 (box.F90:62-63)
    return co, ca, g


def _aml(**kw):
    a = dict(D0up=4.0, Dmdown=3.0, Dup0=2.5, xlamda=35.0, fspco=80.0, Aup00=0.011, Adown10=-0.004, Adown11=0.006,
             hmamin=100.0, hmadmp=0.05, xcexp=1.0, T1_rel=5.0, dT_12=20.0, rho_cp=1.0e3, at2d=5.0e4, at4d=3.0e14,
             ahmd=2.0e4, aface=[-0.01, 0.002], bface=0.003, dface=-0.02)
    a.update(kw)
    return a


def test_load_grid():
    co, ca, g = _grid()
    assert (g.ndxr, g.nxtoar, g.nytoar, g.nxpoar, g.nypoar) == (8, 4, 4, 5, 5)
    assert (g.nxaooc, g.nyaooc, g.nx1, g.ny1) == (4, 4, 1, 3)
    # an odd margin: the ocean sits one row nearer the south wall (integer division)
    g2 = cr.cpl(_ocean(), _atmos(9, 10))
    assert (g2.ny1, g2.nytoar) == (3, 4)
    g3 = cr.cpl(_ocean(41, 10e3), _atmos(9, 10, 50e3))
    assert (g3.ndxr, g3.nxtoar, g3.nxaooc, g3.ny1) == (5, 8, 8, 1)


def test_derivative_stencils_exact_on_polynomials_and_modes():
    box = _ocean(17)
    dx = box["dx"]
    nyp, nxp = 17, 17
    X = np.arange(nxp)[None, :] * dx + np.zeros((nyp, 1))
    Y = np.arange(nyp)[:, None] * dx + np.zeros((1, nxp))
    for alpha in (0.0, 0.5, 3.0):
        d = cr.deriv_2bc("dPdx_2bc", box, alpha, 2.0 * X)
        assert np.allclose(d[1:-1, 1:-1], 2.0, rtol=1e-13)
        assert np.allclose(d[1:-1, 0], 2.0 / (0.5 * alpha + 1.0), rtol=1e-13)
        assert np.allclose(d[1:-1, -1], 2.0 / (0.5 * alpha + 1.0), rtol=1e-13)
        assert np.all(d[0] == 0) and np.all(d[-1] == 0)
        d = cr.deriv_2bc("dPdy_2bc", box, alpha, -3.0 * Y)
        assert np.allclose(d[1:-1, 1:-1], -3.0, rtol=1e-13)
        assert np.allclose(d[0, 1:-1], -3.0 / (0.5 * alpha + 1.0), rtol=1e-13)
        assert np.all(d[:, 0] == 0) and np.all(d[:, -1] == 0)
        # second derivatives: 2a inside; at a wall the mixed condition alpha/(alpha/2+1) (p2-p1)/dx^2
        q = 0.5 * X ** 2 + 3.0 * X
        d = cr.deriv_2bc("dP2dx2_bc", box, alpha, q)
        assert np.allclose(d[1:-1, 1:-1], 1.0, rtol=1e-9)
        b1 = alpha / (0.5 * alpha + 1.0) * (q[:, 1] - q[:, 0]) / dx ** 2
        assert np.allclose(d[1:-1, 0], b1[1:-1], rtol=1e-12, atol=1e-30)
        bn = -alpha / (0.5 * alpha + 1.0) * (q[:, -1] - q[:, -2]) / dx ** 2
        assert np.allclose(d[1:-1, -1], bn[1:-1], rtol=1e-12, atol=1e-30)
        q = -0.25 * Y ** 2
        d = cr.deriv_2bc("dP2dy2_bc", box, alpha, q)
        assert np.allclose(d[1:-1, 1:-1], -0.5, rtol=1e-9)
        assert np.allclose(d[0, 1:-1], alpha / (0.5 * alpha + 1.0) * (q[1, 1:-1] - q[0, 1:-1]) / dx ** 2, rtol=1e-12,
                           atol=1e-30)
    # channel: a discrete Fourier mode in x is an eigenfunction of both x stencils, the wrap included
    ch = _atmos(17, 9)
    dxa = ch["dx"]
    k = 2
    th = 2 * np.pi * k * np.arange(17) / 16
    p = np.cos(th)[None, :] + np.zeros((9, 1))
    d = cr.deriv_2bc("dPdx_2bc", ch, 1.0, p)
    exp = -np.sin(2 * np.pi * k / 16) / dxa * np.sin(th)
    assert np.abs(d[1:-1] - exp[None, :]).max() <= 1e-12 / dxa
    d = cr.deriv_2bc("dP2dx2_bc", ch, 1.0, p)
    assert np.abs(d[1:-1] - 2 * (np.cos(2 * np.pi * k / 16) - 1) / dxa ** 2 * p[1:-1]).max() <= 1e-12 / dxa ** 2


def test_bicubic_reproduces_bicubic_polynomials():
    """Given exact values and (cell-scaled) derivatives at the coarse points, the
    interpolant of any polynomial of degree <= 3 in each variable is exact at every
    fine point: this fixes the Hermite matrix, the order of the corner data and where
    each cell's block lands on the fine grid."""
    co, ca, g = _grid()
    n = g.ndxr
    st = cr.bcuini(n)
    rng = np.random.default_rng(1)
    c = rng.standard_normal((4, 4))
    X = np.arange(g.nxpoar)[None, :] + np.zeros((g.nypoar, 1))
    Y = np.arange(g.nypoar)[:, None] + np.zeros((1, g.nxpoar))

    def poly(x, y, dxo=0, dyo=0):
        s = 0.0
        for i in range(4):
            for j in range(4):
                if i < dxo or j < dyo:
                    continue
                ci = np.prod(np.arange(i, i - dxo, -1)) if dxo else 1.0
                cj = np.prod(np.arange(j, j - dyo, -1)) if dyo else 1.0
                s = s + c[i, j] * ci * cj * x ** (i - dxo) * y ** (j - dyo)
        return s

    out = cr.new_bicubic(g, 33, 33, st, poly(X, Y), poly(X, Y, 1), poly(X, Y, 0, 1), poly(X, Y, 1, 1))
    xf = np.arange(33)[None, :] / n + np.zeros((33, 1))
    yf = np.arange(33)[:, None] / n + np.zeros((1, 33))
    exact = poly(xf, yf)
    assert np.abs(out - exact).max() <= 1e-12 * np.abs(exact).max()


def _stress_invariants(tx, ty, u, v, cd, k):
    """The drag law the stress formula solves: |u|^2 = (|tau|/cd)(1 + k^2 |tau|/cd), and the
    stress turned from u by atan(k sqrt(|tau|/cd)) in the sense of k."""
    t = np.hypot(tx, ty)
    s = t / cd
    lhs, rhs = u * u + v * v, s * (1.0 + k * k * s)
    assert np.abs(lhs - rhs).max() <= 1e-10 * lhs.max()
    cross, dot = ty * u - tx * v, tx * u + ty * v
    m = t > 1e-6 * t.max()
    assert np.allclose(cross[m] / dot[m], k * np.sqrt(s[m]), rtol=1e-9)


@pytest.mark.parametrize("tau_udiff", [False, True])
def test_coupled_stress_drag_law(tau_udiff):
    """p_a quadratic in y: the atmospheric geostrophic velocity is linear in y, the
    derivative stencils and the bicubic interpolant are exact for it inside the
    channel, so the stress on the ocean grid satisfies the drag law with the closed-form
    velocity (minus the ocean's when the stress depends on the velocity difference)."""
    co, ca, g = _grid(tau_udiff)
    nyo, nxo, nya, nxa = 33, 33, 9, 9
    a2 = 40.0 / (ca["dx"] * 8) ** 2
    ya = np.arange(nya)[:, None] * ca["dx"]
    pma = a2 * (ya - 4.0 * ca["dx"]) ** 2 + np.zeros((1, nxa))
    yo = g.yo0 + np.arange(nyo)[:, None] * co["dx"] + np.zeros((1, nxo))
    go_ = 0.3 / (co["dx"] * 32)
    pmo = go_ * (np.arange(nxo)[None, :] + 0.5 * np.arange(nyo)[:, None]) * co["dx"]
    tauxo, tauyo, tauxa, tauya = cr.xforc(co, ca, g, pmo, pma)
    ua = -(2 * a2 * (yo - 4.0 * ca["dx"])) / F0
    va = np.zeros_like(ua)
    raoro = g.rho_a / g.rho_o
    if tau_udiff:
        # the ocean velocity from its pm (dPdx_2bc, zero on the walls it does not reach here)
        dpdx = cr.deriv_2bc("dPdx_2bc", co, co["bcco"], pmo)
        dpdy = cr.deriv_2bc("dPdy_2bc", co, co["bcco"], pmo)
        ua = ua + dpdy / F0
        va = va - dpdx / F0
        k = (g.cdat / F0) * (1.0 / g.hma + raoro / g.hmo)
    else:
        k = (g.cdat / F0) / g.hma
    _stress_invariants(tauxo / raoro, tauyo / raoro, ua, va, g.cdat, k)
    # the atmosphere's own stress from its own velocity (windstress.F90:129-145)
    ka = (g.cdat / F0) / g.hma
    dpx = cr.deriv_2bc("dPdx_2bc", ca, ca["bcco"], pma)
    dpy = cr.deriv_2bc("dPdy_2bc", ca, ca["bcco"], pma)
    if not tau_udiff:
        _stress_invariants(tauxa, tauya, -dpy / F0, dpx / F0, g.cdat, ka)
    else:
        # over the ocean, the atmosphere takes the ocean-resolution stress at the coincident points
        for ja in range(g.ny1, g.ny1 + g.nypoar):
            for ia in range(g.nx1, g.nx1 + g.nxpoar):
                o = ((ja - g.ny1) * g.ndxr, (ia - g.nx1) * g.ndxr)
                assert abs(tauxa[ja - 1, ia - 1] - tauxo[o] / raoro) <= 1e-14 * abs(tauxo).max() / raoro
                assert abs(tauya[ja - 1, ia - 1] - tauyo[o] / raoro) <= 1e-14 * abs(tauyo).max() / raoro


def test_bilint_exact_on_linear_fields():
    co, ca, g = _grid()
    nyt_a, nxt_a = 8, 8
    yta = g.ya0 + (np.arange(nyt_a) + 0.5) * ca["dx"]
    xta = g.xa0 + (np.arange(nxt_a) + 0.5) * ca["dx"]
    yto = g.yo0 + (np.arange(32) + 0.5) * co["dx"]
    xto = g.xo0 + (np.arange(32) + 0.5) * co["dx"]
    f = 2.0 + 1e-5 * yta[:, None] + np.zeros((1, nxt_a))
    out = cr.bilint(ca, co, g, f)
    assert np.allclose(out, (2.0 + 1e-5 * yto)[:, None] + 0 * xto[None, :], rtol=1e-14)
    # linear in x, away from the periodic seam; across it the interpolant joins column nxt to column 1
    h = 1e-5 * xta[None, :] + np.zeros((nyt_a, 1))
    out = cr.bilint(ca, co, g, h)
    inner = (xto >= xta[0]) & (xto <= xta[-1])
    assert np.allclose(out[:, inner], 1e-5 * xto[inner][None, :], rtol=1e-13)
    seam = xto < xta[0]
    w = (xto[seam] - (xta[0] - ca["dx"])) / ca["dx"]
    assert np.allclose(out[:, seam], ((1 - w) * h[0, -1] + w * h[0, 0])[None, :], rtol=1e-13)


def test_compute_forcing_uniform_closed_form_and_conservation():
    co, ca, g = _grid()
    a = _aml(fspco=0.0)
    S, A = 1.5, -2.0
    nyo, nxo, nya, nxa = 32, 32, 8, 8
    z = np.zeros
    fo, fa = cr.compute_forcing(co, ca, g, a, np.full((nyo, nxo), S), np.full((nya, nxa), A),
                                np.full((nya, nxa), g.hma), z((9, 9)), z((9, 9)))
    assert np.allclose(fo, -a["Dmdown"] * A - a["D0up"] * S - a["xlamda"] * (S - A), rtol=1e-14)
    over = np.zeros((nya, nxa), bool)
    over[g.ny1 - 1:g.ny1 - 1 + g.nyaooc, g.nx1 - 1:g.nx1 - 1 + g.nxaooc] = True
    assert np.allclose(fa[~over], -a["Dup0"] * A, rtol=1e-14)
    exp = -a["Dup0"] * A + (a["Dmdown"] - a["Dup0"]) * A + (a["D0up"] * S + a["xlamda"] * (S - A))
    assert np.allclose(fa[over], exp, rtol=1e-13)
    # what leaves the ocean grid reaches the atmosphere's, area for area
    rng = np.random.default_rng(2)
    sst, ast = rng.standard_normal((nyo, nxo)), rng.standard_normal((nya, nxa))
    hm = g.hma + 50 * rng.standard_normal((nya, nxa))
    eta, dtop = rng.standard_normal((9, 9)), rng.standard_normal((9, 9))
    _, f1 = cr.compute_forcing(co, ca, g, a, sst, ast, hm, eta, dtop)
    _, f0 = cr.compute_forcing(co, ca, g, dict(a, D0up=0.0, xlamda=0.0), sst, ast, hm, eta, dtop)
    asto = cr.bilint(ca, co, g, ast)
    left = ((a["D0up"] * sst + a["xlamda"] * (sst - asto)).sum()) * co["dx"] ** 2
    # D0up and xlamda enter fnetat only through the flux gathered from the ocean cells
    got = (f1 - f0).sum() * ca["dx"] ** 2
    assert abs(got - left) <= 1e-12 * np.abs((a["D0up"] * sst + a["xlamda"] * (sst - asto))).sum() * co["dx"] ** 2
    # the thickness and topography terms at T points are the cell means of the p fields
    a3 = _aml(fspco=0.0, D0up=0.0, xlamda=0.0, Dmdown=0.0, Dup0=0.0, hmadmp=0.0)
    Y = np.arange(9)[:, None] + 0.0 * np.arange(9)[None, :]
    _, fe = cr.compute_forcing(co, ca, g, a3, z((nyo, nxo)), z((nya, nxa)), np.full((nya, nxa), g.hma), Y, z((9, 9)))
    assert np.allclose(fe, -a3["Adown11"] * (np.arange(8)[:, None] + 0.5) + 0 * fe, rtol=1e-13)


def test_fsprim_has_zero_mean_over_the_atmosphere():
    co, ca, g = _grid()
    a = _aml(D0up=0.0, Dmdown=0.0, Dup0=0.0, xlamda=0.0, Aup00=0.0, Adown10=0.0, Adown11=0.0, hmadmp=0.0)
    z = np.zeros
    _, fa = cr.compute_forcing(co, ca, g, a, z((32, 32)), z((8, 8)), np.full((8, 8), g.hma), z((9, 9)), z((9, 9)))
    yrel = (np.arange(8) + 0.5) * ca["dx"] - 4 * ca["dx"]
    assert np.allclose(fa[:, 0], -a["fspco"] * 0.5 * np.sin(np.pi * yrel / (8 * ca["dx"])), rtol=1e-13)
    assert abs(fa.sum()) <= 1e-12 * np.abs(fa).sum()
    # the ocean's rows are measured from the ocean's own centre (with the atmosphere's yl)
    fo, _ = cr.compute_forcing(co, ca, g, a, z((32, 32)), z((8, 8)), np.full((8, 8), g.hma), z((9, 9)), z((9, 9)))
    yro = (np.arange(32) + 0.5) * co["dx"] - 16 * co["dx"]
    assert np.allclose(fo[:, 0], -a["fspco"] * 0.5 * np.sin(np.pi * yro / (8 * ca["dx"])), rtol=1e-13)
    assert abs(fo.sum()) <= 1e-12 * np.abs(fo).sum()


def _aml_case(**kw):
    co, ca, g = _grid()
    a = _aml(**kw)
    z = np.zeros
    return ca, g, a, z((9, 9)), z((9, 9)), z((9, 9)), z((2, 9, 9)), z((9, 9))


def test_aml_uniform_states():
    """No flow: (i) stable column (T1_rel - astm >= 2 hmadmp tdta/rho_cp): h relaxes by
    the diabatic term, clipped at hmamin with the matching temperature fix; (ii) the
    other branch resets h to hm; convection caps ast at T1_rel; the entrainment is the
    closed form of its three T-point terms, plus the eta and topography terms at p points."""
    tdta = 1200.0
    F = 30.0
    for astm0, hmix0, hmamin in ((-4.0, 900.0, 100.0), (-4.0, 1500.0, 1499.5), (4.999, 1500.0, 100.0)):
        ca, g, a, p1, tx, ty, eta, dtop = _aml_case(ahmd=0.0, hmamin=hmamin)
        hm = g.hma
        ast0 = astm0 + 0.1
        fn = np.full((8, 8), F)
        xc1 = np.full((8, 8), 0.7)
        res = cr.aml_step(ca, g, a, p1, tx, ty, eta, dtop, fn, xc1, tdta, np.full((8, 8), ast0), np.full((8, 8), astm0),
                          np.full((8, 8), hmix0 + 5.0), np.full((8, 8), hmix0))
        ast, astm, hmixa, hmixam, ent = res
        dT = a["T1_rel"] - astm0
        hdr = a["hmadmp"] / a["rho_cp"]
        if dT >= 2 * hdr * tdta:
            hnew = hmix0 + tdta * (0.0 - hdr * (hmix0 - hm) / dT)
            fix = max(a["hmamin"] - hnew, 0.0)
            hnew, dTfix = hnew + fix, fix * dT / hmix0
        else:
            hnew, dTfix = hm, 0.0
        an = astm0 + tdta * (F / a["rho_cp"] / hmix0) + dTfix
        conv = min(0.0, a["T1_rel"] - an)
        assert np.allclose(hmixa, hnew, rtol=1e-12) and np.all(hmixam == hmix0 + 5.0)
        assert np.allclose(ast, an + conv, rtol=1e-12) and np.all(astm == ast0)
        xfa = a["xcexp"] * a["bface"] * (hmix0 - hm) + a["dface"] * (a["xcexp"] * astm0 + (1 - a["xcexp"]) * 0.7) + \
            a["xcexp"] / (tdta * a["dT_12"]) * (hmix0 + 5.0) * conv
        assert np.allclose(ent, xfa, rtol=1e-12, atol=1e-15)
    # the eta and topography terms (p points)
    ca, g, a, p1, tx, ty, eta, dtop = _aml_case(xcexp=0.0, ahmd=0.0)
    rng = np.random.default_rng(4)
    eta = rng.standard_normal((2, 9, 9))
    dtop = rng.standard_normal((9, 9))
    xc1 = np.zeros((8, 8))
    *_, ent = cr.aml_step(ca, g, a, p1, tx, ty, eta, dtop, np.zeros((8, 8)), xc1, tdta, np.zeros((8, 8)),
                          np.zeros((8, 8)), np.full((8, 8), g.hma), np.full((8, 8), g.hma))
    exp = a["aface"][0] * eta[0] + a["aface"][1] * eta[1] + a["bface"] * dtop
    assert np.allclose(ent, exp, rtol=1e-12, atol=1e-15)


def test_aml_mixed_layer_depth_walls_and_diffusion():
    """hmixa diffuses with ahmd between flux walls held at hm (init_atmos_ml, amlsubs.F90:86-87):
    a uniform excess d over hm decays only in the outer rows, by -ahmd d/dx^2 per unit time;
    the temperature has no flux walls, so a uniform ast has no diffusive tendency."""
    ca, g, a, p1, tx, ty, eta, dtop = _aml_case(hmadmp=0.0, at2d=7e4, at4d=1e15)
    tdta, d = 600.0, 40.0
    hm = g.hma
    res = cr.aml_step(ca, g, a, p1, tx, ty, eta, dtop, np.zeros((8, 8)), np.zeros((8, 8)), tdta, np.full((8, 8), -1.0),
                      np.full((8, 8), -1.0), np.full((8, 8), hm + d), np.full((8, 8), hm + d))
    ast, _, hmixa, _, _ = res
    dx = ca["dx"]
    exp = np.full((8, 8), hm + d)
    exp[0] += tdta * (a["ahmd"] * (-d / dx ** 2))
    exp[-1] += tdta * (a["ahmd"] * (-d / dx ** 2))
    assert np.allclose(hmixa, exp, rtol=1e-13)
    assert np.allclose(ast, -1.0, rtol=1e-14)
