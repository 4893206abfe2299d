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
