"""Pins of the numpy restatement of the diagnostics (oracle/monitor.py) by
closed forms: exact trapezoid integrals, the boundary factors of the _2bc
derivatives, the kinetic energy and CFL maxima of fields linear in y, and the
interface displacements of layers offset by constants."""
import numpy as np

from oracle import monitor as M

NX, NY, DX, DY, F0, ALPHA = 13, 9, 3.0e4, 2.0e4, 1.0e-4, 0.5


def _grid():
    x = np.arange(NX) * DX
    y = np.arange(NY) * DY
    return np.meshgrid(x, y)


def test_trapezoid_integrals_are_exact_for_linear_fields():
    X, Y = _grid()
    xl, yl = (NX - 1) * DX, (NY - 1) * DY
    assert np.isclose(M.int_P_dA(np.ones((NY, NX)), DX, DY), xl * yl, rtol=1e-14)
    assert np.isclose(M.int_P_dA(X, DX, DY), 0.5 * xl * xl * yl, rtol=1e-14)
    assert np.isclose(M.int_P_dA(3.0 * Y + 2.0 * X, DX, DY), 1.5 * yl * yl * xl + xl * xl * yl, rtol=1e-14)
    assert np.isclose(M.avg_P(np.full((NY, NX), 7.0), DX, DY), 7.0, rtol=1e-15)


def test_boundary_derivatives():
    X, Y = _grid()
    a = 2.5
    for cyclic in (False, True):
        dx = M.dPdx_2bc(a * X, DX, ALPHA, cyclic)
        assert np.allclose(dx[1:-1, 1:-1], a, rtol=1e-14)
        assert np.all(dx[0] == 0) and np.all(dx[-1] == 0)
        dy = M.dPdy_2bc(a * Y, DY, ALPHA, cyclic)
        assert np.allclose(dy[1:-1, 1:-1], a, rtol=1e-14)
        assert np.allclose(dy[0, 1:-1], a / (0.5 * ALPHA + 1.0), rtol=1e-14)
    dx = M.dPdx_2bc(a * X, DX, ALPHA, False)
    assert np.allclose(dx[1:-1, 0], a / (0.5 * ALPHA + 1.0), rtol=1e-14)
    assert np.all(M.dPdy_2bc(a * Y, DY, ALPHA, False)[:, 0] == 0)


def test_diagnostics_of_linear_layers():
    X, Y = _grid()
    a, c, h, rho, dt = 3.0e-3, 0.2, [300.0, 700.0], 1000.0, 1800.0
    p = np.stack([a * Y, a * Y - c])
    q = np.stack([np.full((NY, NX), 1e-5), np.full((NY, NX), -2e-5)])
    d = M.diagnose(q, p, DX, DY, F0, ALPHA, False, h, [0.02], -1, rho, dt)
    # u = -a/f0 inside, -a/((0.5 alpha + 1) f0) on the N/S rows, 0 on the E/W columns (box); v = 0
    u_in, u_ns = a / F0, a / ((0.5 * ALPHA + 1.0) * F0)
    ke = ((NY - 2) * (NX - 2) * u_in ** 2 + 2 * 0.5 * (NX - 2) * u_ns ** 2) * DX * DY
    assert np.allclose(d["ekei"], 0.5 * rho * np.array(h) * ke, rtol=1e-13)
    assert np.allclose(d["umax"], a / F0, rtol=1e-14) and np.all(d["vmax"] == 0)
    assert np.allclose(d["cflu"], a / F0 * dt / DX, rtol=1e-14)
    assert np.allclose(d["qavg"], [1e-5, -2e-5], rtol=1e-14)
    assert np.allclose(d["pavg"], [a * (NY - 1) * DY / 2, a * (NY - 1) * DY / 2 - c], rtol=1e-14)
    eta = -1 * c / 0.02
    assert np.isclose(d["etam"][0], eta, rtol=1e-14) and np.isclose(d["et2m"][0], eta * eta, rtol=1e-14)
    assert np.allclose(d["circ"], 1e-6 * (0.0 - a * (NY - 1) * DY) / F0, rtol=1e-14)
    assert d["nonfinite"] == 0.0
    p2 = p.copy()
    p2[1, 3, 4] = np.nan
    q2 = q.copy()
    q2[0, 0, 0] = np.inf
    assert M.diagnose(q2, p2, DX, DY, F0, ALPHA, False, h, [0.02], -1, rho, dt)["nonfinite"] == 2.0


def test_check_continuity_known_answers():
    # continuity holds: dz_sign (int_p_k - int_p_k+1) == dpi_k
    e, f = M.check_continuity([3.0, 1.0, -2.0], [2.0, 3.0], [0.02, 0.01], 1, 1e5, 1e5, 2400.0)
    assert list(e) == [0.0, 0.0] and list(f) == [0.0, 0.0]
    # the other sign convention: est1 = -(2, 3) against dpi = (2, 3)
    e, f = M.check_continuity([3.0, 1.0, -2.0], [2.0, 3.0], [0.02, 0.01], -1, 1e5, 1e5, 2400.0)
    assert list(e) == [-4.0, -6.0] and list(f) == [-2.0, -2.0]
    # below 1e-13 xl yl tdt gp the fraction is reported as 0, the error is not
    # (threshold 1e-13 * 1e3 * 1e3 * 100 * 0.02 = 2e-7)
    e, f = M.check_continuity([0.0, 0.0], [1e-8], [0.02], 1, 1e3, 1e3, 100.0)
    assert e[0] == -1e-8 and f[0] == 0.0
    e, f = M.check_continuity([0.0, 0.0], [1e-6], [0.02], 1, 1e3, 1e3, 100.0)
    assert e[0] == -1e-6 and f[0] == -2.0


# ---------------------------------------------------------------- energetics

def test_del2_restatement_equals_the_step_oracle():
    """monitor.del2_P_bc against the step oracle's del2_P_bc (oracle/cpuref.c,
    pinned by quadratics and the boundary-condition limits), box and channel."""
    from oracle import cpuref as cr

    rng = np.random.default_rng(21)
    for cyclic in (False, True):
        nxp, nyp, dx = 17, 11, 1.5e4
        cfg = dict(nxp=nxp, nyp=nyp, nl=3, cyclic=cyclic, dx=dx, dy=dx, fnot=1e-4, beta=2e-11,
                   h=[300.0, 1100.0, 2600.0], gp=[0.05, 0.025], ah2=[0.0] * 3, ah4=[2e10] * 3, bcco=0.7,
                   delek=2.0, dz_sign=-1, tau_sign=1)
        f = rng.standard_normal((3, nyp, nxp))
        if cyclic:
            f[..., -1] = f[..., 0]
        ref = cr.Basin(cfg).del2(f)
        for k in range(3):
            d = M.del2_P_bc(f[k], dx, dx, 0.7, cyclic)
            assert np.abs(d - ref[k]).max() <= 1e-14 * np.abs(ref[k]).max(), (cyclic, k)


PHASE = 0.3


def _channel_wave(nxp, nyp, m, amp):
    """p = amp cos(k x + PHASE) in every row, k = 2 pi m / xl (periodic in x)."""
    x = np.arange(nxp) * DX
    kx = 2.0 * np.pi * m / ((nxp - 1) * DX)
    return np.broadcast_to(amp * np.cos(kx * x + PHASE), (nyp, nxp)).copy(), kx


def test_channel_dissipation_closed_form():
    """ah2d and ah4d of p = pm = A cos(kx + phi) in a channel.  ug = ugeos = 0;
    the lagged and current v are V_i = -A sin(k x_i + phi) sin(k dx)/(dx f0) on rows
    1..nyp-2 and 0 on the N/S rows.  The x second difference of a periodic
    sinusoid is lambda = 2 (cos(k dx) - 1)/dx^2 times it, so del2(vg) = c_j V_i
    with c_0 = c_{n-1} = fy a/(a/2+1), c_1 = c_{n-2} = lambda - fy and lambda in
    between (fy = 1/dy^2), del4 = (lambda c_j + fy (c_{j-1} + c_{j+1} - 2 c_j))
    V_i on rows 1..nyp-2, and the trapezoid sum of sin^2 over a period is
    (nxp-1)/2."""
    nxp, nyp, m, amp = 25, 9, 3, 40.0
    p1, kx = _channel_wave(nxp, nyp, m, amp)
    p = np.stack([p1, 0.5 * p1])
    h, ah2, ah4, rho = [500.0, 1500.0], [300.0, 200.0], [1e10, 3e10], 1000.0
    e = M.energetics(p, p, np.zeros((nyp, nxp)), np.zeros((nyp, nxp)), np.zeros((nyp, nxp)),
                     np.zeros((nyp, nxp)), DX, DY, F0, ALPHA, True, h, [0.02], ah2, ah4, 1.0, 1, -1, rho)
    lam = 2.0 * (np.cos(kx * DX) - 1.0) / DX ** 2
    fy = 1.0 / DY ** 2
    c = np.full(nyp, lam)
    c[0] = c[-1] = fy * ALPHA / (0.5 * ALPHA + 1.0)
    c[1] = c[-2] = lam - fy
    s2 = c[1:-1].sum()
    s4 = sum(lam * c[j] + fy * (c[j - 1] + c[j + 1] - 2.0 * c[j]) for j in range(1, nyp - 1))
    for k, scale in enumerate((1.0, 0.5)):
        S = (scale * amp * np.sin(kx * DX) / (DX * F0)) ** 2 * (nxp - 1) / 2.0
        assert np.isclose(e["ah2d"][k], -rho * ah2[k] * h[k] * DX * DY * S * s2, rtol=1e-11)
        assert np.isclose(e["ah4d"][k], rho * ah4[k] * h[k] * DX * DY * S * s4, rtol=1e-11)
    # no zonal flow anywhere: the jet value is 0 (first row wins), the drag is the v part
    assert np.all(e["val"] == 0.0) and np.all(e["pos"] == 1)
    S1 = (0.5 * amp * np.sin(kx * DX) / (DX * F0)) ** 2 * (nxp - 1) / 2.0
    assert np.isclose(e["btdg"], 0.5 * rho * 1.0 * F0 * DX * DY * S1 * (nyp - 2), rtol=1e-12)


def test_wind_work_drag_jet_and_forcing_averages_of_linear_layers():
    """p = pm = a y - c_k in a box: ugeos = ug = -a/f0 inside, -a/((alpha/2+1) f0)
    on the N/S rows, 0 on the E/W columns; v = 0.  Uniform stress, pumping and
    entrainment give closed-form utau, btdg, jet, averages and pken."""
    X, Y = _grid()
    a, rho, delek, gp = 3.0e-3, 1000.0, 2.0, [0.02, 0.01]
    cs = [0.0, 0.2, 0.5]
    p = np.stack([a * Y - ck for ck in cs])
    tx, ty, wk, en = 1.5e-4, -2e-5, -3e-6, 4e-7
    one = np.ones((NY, NX))
    e = M.energetics(p, p, wk * one, en * one, tx * one, ty * one, DX, DY, F0, ALPHA, False,
                     [300.0, 700.0, 1200.0], gp, [0.0] * 3, [0.0] * 3, delek, 2, -1, rho)
    u_in, u_ns = -a / F0, -a / ((0.5 * ALPHA + 1.0) * F0)
    int_u = ((NY - 2) * (NX - 2) * u_in + 2 * 0.5 * (NX - 2) * u_ns) * DX * DY
    assert np.isclose(e["utau"], rho * tx * int_u, rtol=1e-13)
    int_u2 = ((NY - 2) * (NX - 2) * u_in ** 2 + 2 * 0.5 * (NX - 2) * u_ns ** 2) * DX * DY
    assert np.isclose(e["btdg"], 0.5 * rho * delek * F0 * int_u2, rtol=1e-13)
    xl = (NX - 1) * DX
    assert np.allclose(e["val"], abs(u_in) * (NX - 2) * DX / xl, rtol=1e-13)
    assert np.isclose(e["wepm"], wk, rtol=1e-14) and np.isclose(e["wapm"], -wk, rtol=1e-14)
    assert np.isclose(e["entm"], en, rtol=1e-14) and np.isclose(e["enam"], en, rtol=1e-14)
    eta = -1 * (cs[1] - cs[0]) / gp[0]   # dz_sign (p_1 - p_2)/gp_1 with p_1 - p_2 = c_2 - c_1
    assert np.isclose(e["pken"], rho * gp[0] * eta * en, rtol=1e-13)
    assert np.all(e["ah2d"] == 0.0) and np.all(e["ah4d"] == 0.0)


def test_jet_position_and_value():
    """p = a y^2: the zonal flow -2 a y / f0 grows northward, so the jet is the
    last row of 1..nyt (the north wall row is not among them); a second layer
    with p = -a (y - y_s)^2 peaks on the south row's side instead."""
    X, Y = _grid()
    a = 1.0e-9
    ys = 2 * DY
    p = np.stack([a * Y ** 2, -a * (Y - ys) ** 2])
    z = np.zeros((NY, NX))
    e = M.energetics(p, p, z, z, z, z, DX, DY, F0, ALPHA, False, [300.0, 700.0], [0.02], [0.0] * 2, [0.0] * 2,
                     2.0, 1, -1, 1000.0)
    xl = (NX - 1) * DX
    yn = (NY - 2) * DY
    assert e["pos"][0] == NY - 1
    assert np.isclose(e["val"][0], 2 * a * yn / F0 * (NX - 2) * DX / xl, rtol=1e-12)
    # layer 2: |u| = 2 a |y - ys| / f0 inside is largest on the far (north) interior row
    assert e["pos"][1] == NY - 1
    assert np.isclose(e["val"][1], 2 * a * (yn - ys) / F0 * (NX - 2) * DX / xl, rtol=1e-12)
