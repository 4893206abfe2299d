"""Pins for the oracle functions that the building-block tests leave open.

Each check compares the oracle (oracle/cpuref.c) with something it did not
compute: a closed form of the reference's formulas on fields where the grid
sums are exact (linear and quadratic fields, trigonometric rows), an identity
between two reference routines that must agree where their formulas coincide,
an invariant that the reference's constraint solve is built to satisfy, or a
dense brute-force solve.  Every test names the reference lines it fixes.
tests/test_oracle_mutations.py plants plausible slips in copies of the oracle
and shows that these tests fail on each of them.
"""
import numpy as np
import pytest

from oracle import cpuref as cr

G = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "golden.npz"))
F0, BETA = 1.0e-4, 2.0e-11
H3, GP3 = [300.0, 1100.0, 2600.0], [0.05, 0.025]
A3 = G["eig_ocean3_A"]   # compute_A of these layers, pinned against its golden in test_oracle.py


def _params(nxp, nyp, cyclic=False, dx=10e3, **kw):
    d = dict(nxp=nxp, nyp=nyp, nl=3, cyclic=cyclic, dx=dx, dy=dx, fnot=F0, beta=BETA, h=H3, gp=GP3,
             ah2=[0.0] * 3, ah4=[2e10] * 3, bcco=0.5, delek=2.0, dz_sign=-1, tau_sign=1)
    d.update(kw)
    return d


def _yprel(nyp, dy):
    return (np.arange(nyp) - 0.5 * (nyp - 1)) * dy


def _trap_x(row, dx):
    """Trapezoid rule along a row (the reference's int_P_dx, numerics.F90:206-212)."""
    return (row[..., 1:-1].sum(-1) + 0.5 * (row[..., 0] + row[..., -1])) * dx


def _qcomp_closed(p, nxp, nyp, dx, lap, kt, ddyn, A=A3):
    """q = Lap(p)/f0 + beta*yprel - f0*A p (+ ddyn on layer kt) given the exact Laplacian."""
    q = lap / F0 + BETA * _yprel(nyp, dx)[None, :, None] - F0 * np.einsum("kl,lji->kji", A, p)
    q[kt] += ddyn
    return q


# ------------------------------------------------------------------ qcomp (vorsubs.F90:27-56)

def _quadratic_state(nxp, nyp, dx, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(nxp)[None, :] * dx
    y = np.arange(nyp)[:, None] * dx
    c = rng.standard_normal((3, 6))
    p = np.stack([c[k, 0] * x ** 2 / dx ** 2 + c[k, 1] * y ** 2 / dx ** 2 + c[k, 2] * x * y / dx ** 2
                  + c[k, 3] * x / dx + c[k, 4] * y / dx + c[k, 5] for k in range(3)])
    lap = np.stack([np.full((nyp, nxp), 2.0 * (c[k, 0] + c[k, 1]) / dx ** 2) for k in range(3)])
    return p, lap


@pytest.mark.parametrize("dz_sign", [-1, 1])
def test_qcomp_interior_closed_form_on_quadratics(dz_sign):
    """init_pv / qcomp: on quadratic p the 5-point Laplacian is exact, so interior q is the
    closed form Lap(p)/f0 + beta*yprel - f0*A p, plus ddyn on layer k_topo (vorsubs.F90:42-53)."""
    nxp, nyp, dx = 14, 11, 10e3
    p, lap = _quadratic_state(nxp, nyp, dx, 3)
    ddyn = 1e-6 * np.random.default_rng(4).standard_normal((nyp, nxp))
    prm = _params(nxp, nyp, dx=dx, dz_sign=dz_sign, ddyn=ddyn)
    b = cr.Basin(prm)
    b.set_state(p=p, pm=2.0 * p)
    b.init_state()
    q, qm, _, _ = b.get_state()
    kt = 0 if dz_sign == 1 else 2
    ref = _qcomp_closed(p, nxp, nyp, dx, lap, kt, ddyn)
    refm = _qcomp_closed(2.0 * p, nxp, nyp, dx, 2.0 * lap, kt, ddyn)
    inner = (slice(None), slice(1, -1), slice(1, -1))
    scale = np.abs(ref[inner]).max()
    assert np.abs(q[inner] - ref[inner]).max() <= 1e-12 * scale
    assert np.abs(qm[inner] - refm[inner]).max() <= 1e-12 * scale


# ------------------------------------------------------------------ ocqbdy (vorsubs.F90:58-114)

@pytest.mark.parametrize("cyclic", [False, True])
def test_ocqbdy_equals_qcomp_on_the_ring_and_corner_rule(cyclic):
    """With dx = dy, ocqbdy's difference term fac*(p_in - p_bdy) (vorsubs.F90:76-85) is the mixed
    boundary value of del2_P_bc/f0 (numerics.F90:437-440,459-466), so on every non-corner ring
    point ocqbdy's q equals qcomp's for ANY p.  Corners (box) take the x-difference once and the
    stretching, beta and ddyn terms twice (vorsubs.F90:78-112), i.e. fac*(x-difference) plus twice
    qcomp's corner value, whose del2 part is zero."""
    nxp, nyp, dx = 13, 10, 20e3
    rng = np.random.default_rng(11)
    p = rng.standard_normal((3, nyp, nxp))
    ddyn = 1e-6 * rng.standard_normal((nyp, nxp))
    if cyclic:
        p[:, :, -1] = p[:, :, 0]
        ddyn[:, -1] = ddyn[:, 0]
    b = cr.Basin(_params(nxp, nyp, cyclic, dx=dx, ddyn=ddyn))
    b.set_state(p=p, pm=p)
    b.init_state()
    qq = b.get_state()[0]           # qcomp(p) everywhere
    q0 = np.full((3, nyp, nxp), 123.0)
    b.set_state(q=q0, qm=q0, p=p, pm=p)
    b.ocqbdy()
    q = b.get_state()[0]
    s = np.abs(qq).max()
    # rows j = 1 and nyp, all i (channel) / non-corner i (box)
    ii = slice(None) if cyclic else slice(1, -1)
    for j in (0, nyp - 1):
        assert np.abs(q[:, j, ii] - qq[:, j, ii]).max() <= 1e-13 * s
    assert np.all(q[:, 1:-1, 1:-1] == 123.0)
    if cyclic:
        assert np.all(q[:, 1:-1, 0] == 123.0) and np.all(q[:, 1:-1, -1] == 123.0)
        return
    for i in (0, nxp - 1):
        assert np.abs(q[:, 1:-1, i] - qq[:, 1:-1, i]).max() <= 1e-13 * s
    fac = 0.5 / (0.5 * 0.5 + 1.0) / dx ** 2 / F0     # bcco/(bcco/2+1)/dx^2/f0, bcco = 0.5
    for j, i, i_in in ((0, 0, 1), (0, nxp - 1, nxp - 2), (nyp - 1, 0, 1), (nyp - 1, nxp - 1, nxp - 2)):
        exp = fac * (p[:, j, i_in] - p[:, j, i]) + 2.0 * qq[:, j, i]
        assert np.abs(q[:, j, i] - exp).max() <= 1e-13 * s


# ------------------------------------------------------------------ update_mass_constr (constraint.F90:167-179)

def test_dpi_leapfrog_closed_form_with_uniform_entrainment():
    """update_mass_constr with int(ent1) dA = e*xl*yl (the trapezoid is exact for a constant):
    dpi_{n+1} = dpip_n - tdt*gp_1*E on the first interface only, so after n steps
    dpi_1 = dpi0 - (n/2) c for even n and dpip0 - ((n+1)/2) c for odd n (c = tdt*gp_1*E); the
    second interface alternates between its two initial values.  suppress_comp_mode then averages
    dpi and dpip (driver.F90:181)."""
    n, dx, tdt, e = 17, 10e3, 7200.0, 3.0e-7
    b = cr.Basin(_params(n, n, dx=dx))
    rng = np.random.default_rng(2)
    p = 1e-2 * rng.standard_normal((3, n, n))
    pm = 1e-2 * rng.standard_normal((3, n, n))
    b.set_state(p=p, pm=pm)
    b.init_state()
    c0 = b.get_constr()
    d0, dm0 = c0[0:2].copy(), c0[2:4].copy()
    # init_mass_constr (constraint.F90:145-165): dz_sign*(int p_k - int p_k+1) with the trapezoid rule
    def trap(f):
        return _trap_x(np.stack([_trap_x(f[j], dx) for j in range(n)]), dx)
    for k in range(2):
        assert np.isclose(d0[k], -(trap(p[k]) - trap(p[k + 1])), rtol=1e-12)
        assert np.isclose(dm0[k], -(trap(pm[k]) - trap(pm[k + 1])), rtol=1e-12)
    ent = np.zeros((2, n, n))
    ent[0] = e
    E = e * (n - 1) * dx * (n - 1) * dx
    c = tdt * GP3[0] * E
    for nt in range(1, 8):
        b.step(tdt, ent, np.zeros((n, n)))
        dpi, dpip = b.get_constr()[0:2], b.get_constr()[2:4]
        if nt % 2 == 0:
            e1, e1m = d0[0] - (nt // 2) * c, dm0[0] - (nt // 2) * c
            e2, e2m = d0[1], dm0[1]
        else:
            e1, e1m = dm0[0] - ((nt + 1) // 2) * c, d0[0] - ((nt - 1) // 2) * c
            e2, e2m = dm0[1], d0[1]
        sc = max(abs(d0[0]), abs(dm0[0]), abs(c) * nt)
        assert abs(dpi[0] - e1) <= 1e-13 * sc and abs(dpip[0] - e1m) <= 1e-13 * sc
        assert dpi[1] == e2 and dpip[1] == e2m
    before = b.get_constr()
    b.suppress()
    after = b.get_constr()
    assert np.array_equal(after[0:2], 0.5 * (before[0:2] + before[2:4]))
    assert np.array_equal(after[2:4], before[2:4])


# ------------------------------------------------------------------ init_constraint (constraint.F90:80-143)

def _line_constraints(p, dx, dy, A=A3):
    """The momentum-constraint functionals of init_constraint applied to a state p (nl, nyp, nxp):
    cs = -int(p_2 - p_1)dx/dy + dy f0^2/2 * A int p_1 dx, cn = +int(p_N - p_N-1)dx/dy + dy f0^2/2 * A int p_N dx."""
    ps, pn = _trap_x(p[:, 0, :], dx), _trap_x(p[:, -1, :], dx)
    cs = -_trap_x(p[:, 1, :] - p[:, 0, :], dx) / dy + 0.5 * dy * F0 ** 2 * A @ ps
    cn = _trap_x(p[:, -1, :] - p[:, -2, :], dx) / dy + 0.5 * dy * F0 ** 2 * A @ pn
    return cs, cn


def test_init_constraint_closed_form_for_p_quadratic_in_y():
    """p_k = a_k + b_k y + c_k y^2, independent of x: every line integral is xl times a row value,
    so cs = -xl (P(dy) - P(0))/dy + (dy f0^2/2) xl A P(0) and
    cn = xl (P(yl) - P(yl-dy))/dy + (dy f0^2/2) xl A P(yl), csp/cnp likewise from pm
    (constraint.F90:116-141), with P evaluated from the polynomial.  A is not symmetric here, so a
    transposed A fails; the quadratic term makes each row difference distinct."""
    nxp, nyp, dx = 17, 9, 15e3
    xl, yl = (nxp - 1) * dx, (nyp - 1) * dx
    y = np.arange(nyp)[:, None] * dx
    rng = np.random.default_rng(21)
    cp, cm = rng.standard_normal((3, 3)), rng.standard_normal((3, 3))
    poly = lambda c, yy: c[:, 0] + c[:, 1] * yy / yl + c[:, 2] * (yy / yl) ** 2
    p = np.stack([np.broadcast_to(poly(cp[k:k + 1], y), (nyp, nxp)) for k in range(3)]).copy()
    pm = np.stack([np.broadcast_to(poly(cm[k:k + 1], y), (nyp, nxp)) for k in range(3)]).copy()
    b = cr.Basin(_params(nxp, nyp, True, dx=dx))
    b.set_state(p=p, pm=pm)
    b.init_state()
    con = b.get_constr()
    cs, cn, csp, cnp = con[4:7], con[7:10], con[10:13], con[13:16]
    k = 0.5 * dx * F0 ** 2 * xl

    def south(c):
        return -xl * (poly(c, dx) - poly(c, 0.0)) / dx + k * A3 @ poly(c, 0.0)

    def north(c):
        return xl * (poly(c, yl) - poly(c, yl - dx)) / dx + k * A3 @ poly(c, yl)

    for got, ref in ((cs, south(cp)), (cn, north(cp)), (csp, south(cm)), (cnp, north(cm))):
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(A3 - A3.T).max() > 1e-3 * np.abs(A3).max()


# ------------------------------------------------------------------ ajis / ajin (numerics.F90:374-386)

def test_boundary_jacobian_integrals_closed_form():
    """q = A_j cos(kx) on rows 1, 2, nyp-1, nyp and p = B_j sin(kx) on rows 2 and nyp-1 of a channel
    (k = 2 pi/xl): the centred difference gives dp_dx = 2 B sin(k dx) cos(kx), and the periodic
    trapezoid of cos^2 over N = nxp-1 points is N dx/2, so
    ajis = (A_1 + 2 A_2) B_2 N sin(k dx)/12 and ajin = -(A_nyp + 2 A_nyp-1) B_nyp-1 N sin(k dx)/12."""
    nxp, nyp, dx = 25, 9, 10e3
    N = nxp - 1
    kx = 2 * np.pi / (N * dx)
    x = np.arange(nxp) * dx
    b = cr.Basin(_params(nxp, nyp, True, dx=dx))
    rng = np.random.default_rng(31)
    Aq = rng.standard_normal((3, 4))
    Bp = rng.standard_normal((3, 2))
    q = np.zeros((3, nyp, nxp))
    p = np.zeros((3, nyp, nxp))
    for k in range(3):
        q[k, 0], q[k, 1], q[k, -2], q[k, -1] = (Aq[k, t] * np.cos(kx * x) for t in range(4))
        p[k, 1], p[k, -2] = Bp[k, 0] * np.sin(kx * x), Bp[k, 1] * np.sin(kx * x)
        p[k, 3] = rng.standard_normal() * np.sin(2 * kx * x)    # rows the integrals must not read
        q[k, 3] = rng.standard_normal() * np.cos(kx * x)
    _, ajis, ajin = b.jacobian(p, q)
    S = N * np.sin(kx * dx) / 12.0
    es = (Aq[:, 0] + 2 * Aq[:, 1]) * Bp[:, 0] * S
    en = -(Aq[:, 3] + 2 * Aq[:, 2]) * Bp[:, 1] * S
    assert np.abs(ajis - es).max() <= 1e-13 * np.abs(es).max()
    assert np.abs(ajin - en).max() <= 1e-13 * np.abs(en).max()


# ------------------------------------------------------------------ step_constraints (constraint.F90:182-215)

def test_step_constraints_closed_form_leapfrog():
    """A channel at rest in p with q = beta*y and no friction: every field stays x-independent, so
    the Jacobian and its boundary integrals vanish exactly and only the stress and boundary
    entrainment drive the constraint vectors.  With entrainment uniform along the zonal walls
    (e_s, e_n in layer 1, e2_s, e2_n in layer 2), the right-hand sides are the constants
      R_s = -dy f0/2 [dz f0/h1 e_s xl, dz f0/h1 e2_s xl, 0] + [tau f0/h1 txis, 0, 0]
      R_n = -dy f0/2 [-dz f0/h2 e_n xl, -dz f0/h2 e2_n xl, 0] - [tau f0/h1 txin, 0, 0]
    (vorticity.F90:84-93, with fohfac(2) in the northern layer-1 term) and the leapfrog gives
    cs after n steps = cs0 + (n/2) tdt R_s (n even) or csp0 + ((n+1)/2) tdt R_s (n odd)."""
    nxp, nyp, dx, tdt = 17, 9, 10e3, 7200.0
    xl = (nxp - 1) * dx
    prm = _params(nxp, nyp, True, dx=dx, ah4=[0.0] * 3, delek=0.0, dz_sign=1, tau_sign=-1)
    b = cr.Basin(prm)
    z = np.zeros((3, nyp, nxp))
    b.set_state(p=z, pm=z)
    b.init_state()
    rng = np.random.default_rng(41)
    con0 = b.get_constr()
    con0[4:] = rng.standard_normal(12)
    b.set_constr(con0)
    cs0, cn0, csp0, cnp0 = con0[4:7], con0[7:10], con0[10:13], con0[13:16]
    es, en, e2s, e2n = 1e-6, -2e-6, 3e-7, 5e-7
    txis, txin = 2e-4, -1e-4
    ent = np.zeros((2, nyp, nxp))
    ent[0, 0, :], ent[0, -1, :], ent[1, 0, :], ent[1, -1, :] = es, en, e2s, e2n
    fh = [F0 / h for h in H3]
    Rs = -0.5 * dx * F0 * np.array([fh[0] * es * xl, fh[0] * e2s * xl, 0.0])
    Rn = -0.5 * dx * F0 * np.array([-fh[1] * en * xl, -fh[1] * e2n * xl, 0.0])
    Rs[0] += -1 * fh[0] * txis
    Rn[0] -= -1 * fh[0] * txin
    for nt in range(1, 8):
        b.step(tdt, ent, np.zeros((nyp, nxp)), txis, txin)
        con = b.get_constr()
        cs, cn = con[4:7], con[7:10]
        if nt % 2 == 0:
            ecs, ecn = cs0 + (nt // 2) * tdt * Rs, cn0 + (nt // 2) * tdt * Rn
        else:
            ecs, ecn = csp0 + ((nt + 1) // 2) * tdt * Rs, cnp0 + ((nt + 1) // 2) * tdt * Rn
        sc = max(np.abs(ecs).max(), np.abs(ecn).max())
        assert np.abs(cs - ecs).max() <= 1e-13 * sc and np.abs(cn - ecn).max() <= 1e-13 * sc, nt
        _, _, p, _ = b.get_state()
        assert np.all(p == p[:, :, :1])    # x-independent throughout


# ------------------------------------------------------------------ homsol_cyclic (homog.F90:128-212)

def test_homogeneous_channel_solutions_brute_force():
    """hom_sol_bt = (nyp-j)/(nyp-1); hbsi = yl/xl; hom_sol_bc1/bc2 = L + rdm2*s with
    (d2/dy2 - rdm2) s = L, s = 0 on the walls, L = 1 -> 0 (bc1) or 0 -> 1 (bc2) linearly in y:
    checked against a dense solve of the 1-D operator (the x-independent part of the 5-point one)."""
    nxp, nyp, dx = 17, 9, 10e3
    b = cr.Basin(_params(nxp, nyp, True, dx=dx))
    rdm2 = b.modes()[3]
    bt, b1, b2, coef = b.homog()
    j = np.arange(nyp)
    assert np.allclose(bt, (nyp - 1 - j) / (nyp - 1), rtol=0, atol=1e-15)
    assert np.isclose(coef[-1], (nyp - 1) / (nxp - 1), rtol=1e-15)
    ni = nyp - 2
    D = (np.diag(-2.0 * np.ones(ni)) + np.diag(np.ones(ni - 1), 1) + np.diag(np.ones(ni - 1), -1)) / dx ** 2
    for m in (2, 3):
        for L, got in (((nyp - 1 - j) / (nyp - 1), b1[m - 2]), (j / (nyp - 1), b2[m - 2])):
            s = np.zeros(nyp)
            s[1:-1] = np.linalg.solve(D - rdm2[m - 1] * np.eye(ni), L[1:-1])
            assert np.abs(got - (L + rdm2[m - 1] * s)).max() <= 1e-12


def _modal_functionals(h, rd, dx, dy, xl):
    """South and north constraint functionals of an x-independent modal profile h(y) (homog.F90:160-172)."""
    S = -xl * (h[1] - h[0]) / dy + 0.5 * dy * rd * h[0] * xl
    N = xl * (h[-1] - h[-2]) / dy + 0.5 * dy * rd * h[-1] * xl
    return S, N


def test_channel_homogeneous_coefficients_invert_the_2x2_system():
    """cyclic_homog's coefficients c1 = hc2n Ls - hc2s Ln, c2 = hc1s Ln - hc1n Ls (homog.F90:315-320)
    must make c1*bc1 + c2*bc2 meet any south/north targets (Ls, Ln); and c_bt = Ls*hbsi must meet
    the south target with the barotropic profile."""
    nxp, nyp, dx = 17, 9, 10e3
    xl = (nxp - 1) * dx
    b = cr.Basin(_params(nxp, nyp, True, dx=dx))
    rdm2 = b.modes()[3]
    bt, b1, b2, coef = b.homog()
    for m in (2, 3):
        hc1s, hc2s, hc1n, hc2n = coef[4 * (m - 2): 4 * (m - 2) + 4]
        for Ls, Ln in ((1.0, 0.0), (0.0, 1.0), (0.3, -2.0)):
            c1 = hc2n * Ls - hc2s * Ln
            c2 = hc1s * Ln - hc1n * Ls
            S, N = _modal_functionals(c1 * b1[m - 2] + c2 * b2[m - 2], rdm2[m - 1], dx, dx, xl)
            assert abs(S - Ls) <= 1e-12 and abs(N - Ln) <= 1e-12
    S, _ = _modal_functionals(0.7 * coef[-1] * bt, 0.0, dx, dx, xl)
    assert abs(S - 0.7) <= 1e-13


# ------------------------------------------------------------------ cyclic_homog (homog.F90:280-329)

def _channel_run(nsteps=6, seed=51):
    nxp, nyp, dx, tdt = 33, 17, 10e3, 3600.0
    prm = _params(nxp, nyp, True, dx=dx, ah2=[1e3] * 3, dz_sign=1, tau_sign=-1)
    b = cr.Basin(prm)
    rng = np.random.default_rng(seed)
    p = np.zeros((3, nyp, nxp))
    p[:, 1:-1, :-1] = 1e2 * rng.standard_normal((3, nyp - 2, nxp - 1))
    p[:, 0, :-1] = 50.0 + rng.standard_normal((3, nxp - 1))
    p[:, :, -1] = p[:, :, 0]
    b.set_state(p=p, pm=p)
    b.init_state()
    ent = 1e-6 * rng.standard_normal((2, nyp, nxp))
    ent[:, :, -1] = ent[:, :, 0]
    wekp = 1e-6 * rng.standard_normal((nyp, nxp))
    wekp[:, -1] = wekp[:, 0]
    return b, dx, tdt, ent, wekp, (nxp - 1) * dx


def test_channel_constraint_holds_after_each_step():
    """cyclic_homog picks the homogeneous correction so that the new pressure satisfies the
    leapfrogged momentum constraints: projected on each mode m (with ctl2m, A's left
    eigenvectors), init_constraint's functionals of p_new equal the projected cs for every mode and
    cn for the baroclinic modes (the barotropic coefficient is fixed by the southern boundary only,
    homog.F90:139-148)."""
    b, dx, tdt, ent, wekp, xl = _channel_run()
    _, ctl2m, _, _ = b.modes()
    for nt in range(1, 7):
        b.step(tdt, ent, wekp, 1e-4, -2e-4)
        _, _, p, _ = b.get_state()
        con = b.get_constr()
        cs, cn = con[4:7], con[7:10]
        fs, fn = _line_constraints(p, dx, dx)
        ds, dn = ctl2m.T @ (fs - cs), ctl2m.T @ (fn - cn)
        sc = max(np.abs(ctl2m.T @ cs).max(), np.abs(ctl2m.T @ fs).max())
        assert np.abs(ds).max() <= 1e-11 * sc, (nt, ds)
        assert np.abs(dn[1:]).max() <= 1e-11 * sc, (nt, dn)


# ------------------------------------------------------------------ the checks catch planted slips

def test_planted_mistakes_fail():
    """The functionals and closed forms above are sharp: a sign slip in the south functional, a
    transposed A, or a dropped barotropic correction each move the compared quantity far beyond
    the tolerances used."""
    b, dx, tdt, ent, wekp, xl = _channel_run()
    _, ctl2m, _, _ = b.modes()
    for _ in range(3):
        b.step(tdt, ent, wekp, 1e-4, -2e-4)
    _, _, p, _ = b.get_state()
    con = b.get_constr()
    cs = con[4:7]
    fs, _ = _line_constraints(p, dx, dx)
    sc = np.abs(ctl2m.T @ cs).max()
    assert np.abs(ctl2m.T @ (fs - cs)).max() <= 1e-11 * sc
    ps = _trap_x(p[:, 0, :], dx)
    wrong_sign = _trap_x(p[:, 1, :] - p[:, 0, :], dx) / dx + 0.5 * dx * F0 ** 2 * A3 @ ps
    wrong_A = fs - 0.5 * dx * F0 ** 2 * (A3 - A3.T) @ ps
    no_bt = p.copy()
    no_bt -= p[:, :1, :1]     # a constant shift of every layer changes the A-term only through A @ 1
    for bad in (wrong_sign, wrong_A, _line_constraints(no_bt, dx, dx)[0]):
        assert np.abs(ctl2m.T @ (bad - cs)).max() > 1e-6 * sc


@pytest.mark.parametrize("dz_sign", [-1, 1])
def test_step_constraints_friction_and_drag_terms_closed_form(dz_sign):
    """x-independent pm = g_k(y): del2_P_bc reduces to the 1-D operator D (fy(g_j-1 + g_j+1 - 2g_j)
    inside, the mixed condition +-fy a/(a/2+1)(g_2 - g_1) on the walls), every line integral is xl
    times a row value, and J = 0.  One step must then move cs by
      tdt * (-ap3s + ap5s + delta_{k,nl} (f0/h_nl) bdrins),  ap3s = ah2 xl ((Dg)_2 - (Dg)_1)/dy,
      ap5s = ah4 xl ((DDg)_2 - (DDg)_1)/dy,  bdrins = f0 delek/(2|f0|) xl (g_kt,2 - g_kt,1)/dy
    and cn by tdt * (ap3n - ap5n - delta_{k,nl} (f0/h_nl) bdrinn) (vorticity.F90:57-62,95-101,
    constraint.F90:197-213); csp, cnp take the old cs, cn."""
    nxp, nyp, dx, tdt, bcco = 17, 11, 10e3, 3600.0, 0.5
    xl = (nxp - 1) * dx
    ah2, ah4, delek = [1e3, 2e3, 3e3], [1e9, 2e10, 5e9], 2.0
    prm = _params(nxp, nyp, True, dx=dx, ah2=ah2, ah4=ah4, delek=delek, dz_sign=dz_sign, bcco=bcco)
    b = cr.Basin(prm)
    rng = np.random.default_rng(61)
    g = rng.standard_normal((3, nyp))
    p = np.repeat(g[:, :, None], nxp, axis=2)
    b.set_state(p=p, pm=p)
    b.init_state()
    con0 = b.get_constr()
    con0[4:] = rng.standard_normal(12)
    b.set_constr(con0)
    fy = 1.0 / dx ** 2
    D = np.zeros((nyp, nyp))
    for j in range(1, nyp - 1):
        D[j, j - 1], D[j, j], D[j, j + 1] = fy, -2 * fy, fy
    cb = fy * bcco / (0.5 * bcco + 1.0)
    D[0, 0], D[0, 1], D[-1, -1], D[-1, -2] = -cb, cb, -cb, cb
    d2 = g @ D.T
    d4 = d2 @ D.T
    kt = 0 if dz_sign == 1 else 2
    cd = F0 * delek / (2 * abs(F0))
    Rs = -np.array(ah2) * xl * (d2[:, 1] - d2[:, 0]) / dx + np.array(ah4) * xl * (d4[:, 1] - d4[:, 0]) / dx
    Rn = np.array(ah2) * xl * (d2[:, -1] - d2[:, -2]) / dx - np.array(ah4) * xl * (d4[:, -1] - d4[:, -2]) / dx
    Rs[2] += F0 / H3[2] * cd * xl * (g[kt, 1] - g[kt, 0]) / dx
    Rn[2] -= F0 / H3[2] * cd * xl * (g[kt, -1] - g[kt, -2]) / dx
    b.step(tdt, np.zeros((2, nyp, nxp)), np.zeros((nyp, nxp)), 0.0, 0.0)
    con = b.get_constr()
    cs, cn, csp, cnp = con[4:7], con[7:10], con[10:13], con[13:16]
    ecs, ecn = con0[10:13] + tdt * Rs, con0[13:16] + tdt * Rn
    sc = max(np.abs(ecs).max(), np.abs(ecn).max(), tdt * np.abs(Rs).max())
    assert np.abs(cs - ecs).max() <= 1e-12 * sc and np.abs(cn - ecn).max() <= 1e-12 * sc
    assert np.array_equal(csp, con0[4:7]) and np.array_equal(cnp, con0[7:10])
    assert tdt * np.abs(Rs).max() > 0.1 * sc      # the terms are not negligible
