"""ctypes binding of the CPU oracle (oracle/cpuref.c).

Test infrastructure only: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, never by the
openqg_amd package.  Fields are numpy float64 arrays of shape (nl, nyp, nxp),
the C view of the reference's Fortran (nxp, nyp, nl) arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OQG_CPUREF_LIB points the tests at another build of the oracle (the mutation
# tests load copies with planted mistakes this way).
_LIB_PATH = os.environ.get("OQG_CPUREF_LIB") or os.path.join(_HERE, "libcpuref.so")
MAXL = 8
_c_double_p = ctypes.POINTER(ctypes.c_double)


class CrParams(ctypes.Structure):
    _fields_ = [
        ("nxp", ctypes.c_int), ("nyp", ctypes.c_int), ("nl", ctypes.c_int), ("cyclic", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("fnot", ctypes.c_double), ("beta", ctypes.c_double),
        ("h", ctypes.c_double * MAXL), ("gp", ctypes.c_double * MAXL),
        ("ah2", ctypes.c_double * MAXL), ("ah4", ctypes.c_double * MAXL),
        ("bcco", ctypes.c_double), ("delek", ctypes.c_double),
        ("dz_sign", ctypes.c_int), ("tau_sign", ctypes.c_int), ("k_topo", ctypes.c_int),
    ]


class CrOmlParams(ctypes.Structure):
    _fields_ = [("hm", ctypes.c_double), ("T1_rel", ctypes.c_double), ("dT_12", ctypes.c_double),
                ("rho_cp", ctypes.c_double), ("d2", ctypes.c_double), ("d4", ctypes.c_double),
                ("ycexp", ctypes.c_double), ("nval", ctypes.c_double), ("sval", ctypes.c_double),
                ("fluxn", ctypes.c_int), ("fluxs", ctypes.c_int)]


class CrCpl(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("ndxr", "nxtoar", "nytoar", "nxpoar", "nypoar", "nxaooc", "nyaooc", "nx1",
                                             "ny1")] + \
               [(n, ctypes.c_double) for n in ("xo0", "yo0", "xa0", "ya0", "hmo", "hma", "rho_o", "rho_a", "cdat")] + \
               [("tau_udiff", ctypes.c_int)]


_AML_D = ("D0up", "Dmdown", "Dup0", "xlamda", "fspco", "Aup00", "Adown10", "Adown11", "hmamin", "hmadmp", "xcexp",
          "T1_rel", "dT_12", "rho_cp", "at2d", "at4d", "ahmd")


class CrAmlParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in _AML_D] + [("aface", ctypes.c_double * MAXL), ("bface", ctypes.c_double),
                                                         ("dface", ctypes.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        dp = _c_double_p
        i = ctypes.c_int
        d = ctypes.c_double
        L.cr_create.argtypes = [ctypes.POINTER(CrParams), dp, ctypes.POINTER(vp)]
        L.cr_create.restype = i
        L.cr_destroy.argtypes = [vp]
        L.cr_set_state.argtypes = [vp, dp, dp, dp, dp]
        L.cr_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.cr_constr_len.argtypes = [vp]
        L.cr_constr_len.restype = i
        L.cr_get_constr.argtypes = [vp, dp]
        L.cr_set_constr.argtypes = [vp, dp]
        L.cr_init_state.argtypes = [vp]
        L.cr_step.argtypes = [vp, d, dp, i, dp, d, d]
        L.cr_suppress.argtypes = [vp]
        L.cr_get_modes.argtypes = [vp, dp, dp, dp, dp]
        L.cr_get_homog.argtypes = [vp, dp, dp, dp, dp]
        L.cr_solve_inhomog.argtypes = [vp, i, dp, dp]
        L.cr_del2_bc.argtypes = [vp, dp, dp, i]
        L.cr_jacobian.argtypes = [vp, dp, dp, dp, dp, dp]
        L.cr_int_P_dA.argtypes = [vp, dp]
        L.cr_int_P_dA.restype = d
        L.cr_int_P_dx.argtypes = [vp, dp]
        L.cr_int_P_dx.restype = d
        L.cr_dst1.argtypes = [i, i, dp, dp]
        L.cr_rfft.argtypes = [i, i, dp, dp]
        L.cr_irfft.argtypes = [i, i, dp, dp]
        L.cr_cfft.argtypes = [i, i, i, dp, dp]
        L.cr_ekman_wekp.argtypes = [ctypes.POINTER(CrParams), d, dp, dp, dp, dp, dp]
        L.cr_set_threads.argtypes = [i]
        L.cr_ocqbdy.argtypes = [vp]
        L.cr_thomas.argtypes = [i, d, d, dp, dp]
        PP = ctypes.POINTER(CrParams)
        OP = ctypes.POINTER(CrOmlParams)
        L.cr_del2_T.argtypes = [PP, dp, i, d, i, d, dp]
        L.cr_ekman_uvw.argtypes = [PP, d, dp, dp, dp, dp, dp]
        L.cr_adv_diff.argtypes = [PP, OP, dp, dp, dp, dp, dp, dp]
        L.cr_map_T_to_P.argtypes = [PP, dp, dp]
        L.cr_oml_step.argtypes = [PP, OP, dp, dp, dp, dp, d, i, dp, dp, dp]
        GP = ctypes.POINTER(CrCpl)
        AP = ctypes.POINTER(CrAmlParams)
        L.cr_load_grid.argtypes = [PP, PP, GP]
        L.cr_bcuini.argtypes = [i, dp]
        for n in ("cr_dPdx_2bc", "cr_dPdy_2bc", "cr_dP2dx2_bc", "cr_dP2dy2_bc"):
            getattr(L, n).argtypes = [PP, d, dp, dp]
        L.cr_new_bicubic.argtypes = [GP, i, i, dp, dp, dp, dp, dp, dp]
        L.cr_xforc.argtypes = [PP, PP, GP, dp, dp, dp, dp, dp, dp, dp]
        L.cr_bilint.argtypes = [PP, PP, GP, dp, dp]
        L.cr_compute_forcing.argtypes = [PP, PP, GP, AP, dp, dp, dp, dp, dp, dp, dp]
        L.cr_aml_step.argtypes = [PP, GP, AP, dp, dp, dp, dp, dp, dp, dp, d, i, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_c_double_p)


def _get(cfg, key, default=None):
    if isinstance(cfg, dict):
        return cfg.get(key, default)
    return getattr(cfg, key, default)


def make_params(cfg) -> CrParams:
    nl = int(_get(cfg, "nl"))
    P = CrParams()
    P.nxp, P.nyp, P.nl = int(_get(cfg, "nxp")), int(_get(cfg, "nyp")), nl
    P.cyclic = int(bool(_get(cfg, "cyclic", False)))
    P.dx, P.dy = float(_get(cfg, "dx")), float(_get(cfg, "dy"))
    P.fnot, P.beta = float(_get(cfg, "fnot")), float(_get(cfg, "beta"))
    for name in ("h", "ah2", "ah4"):
        arr = getattr(P, name)
        for k, v in enumerate(_get(cfg, name)):
            arr[k] = float(v)
    for k, v in enumerate(_get(cfg, "gp")):
        P.gp[k] = float(v)
    P.bcco, P.delek = float(_get(cfg, "bcco")), float(_get(cfg, "delek"))
    P.dz_sign, P.tau_sign = int(_get(cfg, "dz_sign")), int(_get(cfg, "tau_sign"))
    kt = _get(cfg, "k_topo", None)
    P.k_topo = int(kt) if kt is not None else (1 if P.dz_sign == 1 else nl)
    return P


def set_threads(n: int):
    lib().cr_set_threads(int(n))


class Basin:
    """One oracle basin: init_qg + state + step_qg."""

    def __init__(self, cfg):
        L = lib()
        self.cfg = cfg
        self.P = make_params(cfg)
        self.nxp, self.nyp, self.nl = self.P.nxp, self.P.nyp, self.P.nl
        ddyn = _get(cfg, "ddyn", None)
        self._ddyn = None if ddyn is None else np.ascontiguousarray(ddyn, dtype=np.float64)
        h = ctypes.c_void_p()
        rc = L.cr_create(ctypes.byref(self.P), _ptr(self._ddyn), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"cr_create failed rc={rc}")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.cr_destroy(h)
            self._h = None

    @property
    def shape(self):
        return (self.nl, self.nyp, self.nxp)

    def _field(self):
        return np.zeros(self.shape, dtype=np.float64)

    def set_state(self, q=None, qm=None, p=None, pm=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        q, qm, p, pm = c(q), c(qm), c(p), c(pm)
        lib().cr_set_state(self._h, _ptr(q), _ptr(qm), _ptr(p), _ptr(pm))

    def get_state(self):
        q, qm, p, pm = (self._field() for _ in range(4))
        lib().cr_get_state(self._h, _ptr(q), _ptr(qm), _ptr(p), _ptr(pm))
        return q, qm, p, pm

    def get_constr(self):
        n = lib().cr_constr_len(self._h)
        out = np.zeros(n)
        lib().cr_get_constr(self._h, _ptr(out))
        return out

    def set_constr(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().cr_set_constr(self._h, _ptr(v))

    def init_state(self):
        lib().cr_init_state(self._h)

    def step(self, tdt, ent, wekp, txis=0.0, txin=0.0):
        ent = np.ascontiguousarray(ent, dtype=np.float64)
        if ent.ndim == 2:
            ent = ent[None]
        wekp = np.ascontiguousarray(wekp, dtype=np.float64)
        lib().cr_step(self._h, float(tdt), _ptr(ent), int(ent.shape[0]), _ptr(wekp), float(txis), float(txin))

    def suppress(self):
        lib().cr_suppress(self._h)

    def ocqbdy(self):
        lib().cr_ocqbdy(self._h)

    def modes(self):
        nl = self.nl
        amat, ctl2m, ctm2l = (np.zeros((nl, nl)) for _ in range(3))
        rdm2 = np.zeros(nl)
        lib().cr_get_modes(self._h, _ptr(amat), _ptr(ctl2m), _ptr(ctm2l), _ptr(rdm2))
        # stored Fortran order (column-major) -> transpose to [row][col] = (i, j)
        return amat.T.copy(), ctl2m.T.copy(), ctm2l.T.copy(), rdm2

    def homog(self):
        nl = self.nl
        if not self.P.cyclic:
            hs = np.zeros((nl - 1, self.nyp, self.nxp))
            lib().cr_get_homog(self._h, _ptr(hs), None, None, None)
            return hs
        bt = np.zeros(self.nyp)
        b1 = np.zeros((nl - 1, self.nyp))
        b2 = np.zeros((nl - 1, self.nyp))
        coef = np.zeros(4 * (nl - 1) + 1)
        lib().cr_get_homog(self._h, _ptr(bt), _ptr(b1), _ptr(b2), _ptr(coef))
        return bt, b1, b2, coef

    def solve_inhomog(self, m, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.zeros((self.nyp, self.nxp))
        lib().cr_solve_inhomog(self._h, int(m), _ptr(rhs), _ptr(out))
        return out

    def start_from_q(self, q0):
        """The initial state by one inversion: q = qm = q0, p = pm = 0 and a zero
        constraint state, then one step with tdt = 0 (q_new = qm, dpi unchanged,
        p by solve_pressure, q's ring by ocqbdy), then qm = q, pm = p."""
        z = self._field()
        self.set_state(q=q0, qm=q0, p=z, pm=z)
        self.set_constr(np.zeros(len(self.get_constr())))
        self.step(0.0, np.zeros((max(self.nl - 1, 1), self.nyp, self.nxp)), np.zeros((self.nyp, self.nxp)))
        q, _, p, _ = self.get_state()
        self.set_state(q=q, qm=q, p=p, pm=p)

    def del2(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        nl = 1 if f.ndim == 2 else f.shape[0]
        out = np.zeros_like(f)
        lib().cr_del2_bc(self._h, _ptr(f), _ptr(out), nl)
        return out

    def jacobian(self, p, q):
        p = np.ascontiguousarray(p, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        jac = np.zeros_like(p)
        ajis = np.zeros(self.nl)
        ajin = np.zeros(self.nl)
        lib().cr_jacobian(self._h, _ptr(p), _ptr(q), _ptr(jac), _ptr(ajis), _ptr(ajin))
        return jac, ajis, ajin

    def int_P_dA(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        return lib().cr_int_P_dA(self._h, _ptr(f))

    def int_P_dx(self, row):
        row = np.ascontiguousarray(row, dtype=np.float64)
        return lib().cr_int_P_dx(self._h, _ptr(row))


def dst1(x):
    """FFTW RODFT00 (unnormalised) along the last axis."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    out = np.zeros_like(x)
    lib().cr_dst1(n, int(x.size // n), _ptr(x), _ptr(out))
    return out


def thomas(a, bb, f):
    """The reference's Numerical Recipes Thomas solve (inhomog.F90:188-199)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.zeros_like(f)
    lib().cr_thomas(int(f.size), float(a), float(bb), _ptr(f), _ptr(u))
    return u


def rfft(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    cnt = int(x.size // n)
    out = np.zeros(x.shape[:-1] + (n // 2 + 1, 2))
    lib().cr_rfft(n, cnt, _ptr(x), _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def irfft(X, n):
    Xr = np.ascontiguousarray(np.stack([X.real, X.imag], axis=-1), dtype=np.float64)
    cnt = int(Xr.size // (2 * (n // 2 + 1)))
    out = np.zeros(X.shape[:-1] + (n,))
    lib().cr_irfft(n, cnt, _ptr(Xr), _ptr(out))
    return out


def cfft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    xr = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1))
    out = np.zeros_like(xr)
    lib().cr_cfft(n, int(sign), int(x.size // n), _ptr(xr), _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def ekman_wekp(cfg, taux, tauy):
    P = make_params(cfg)
    taux = np.ascontiguousarray(taux, dtype=np.float64)
    tauy = np.ascontiguousarray(tauy, dtype=np.float64)
    wekp = np.zeros_like(taux)
    txis = ctypes.c_double()
    txin = ctypes.c_double()
    lib().cr_ekman_wekp(ctypes.byref(P), 0.0, _ptr(taux), _ptr(tauy), _ptr(wekp),
                        ctypes.cast(ctypes.byref(txis), _c_double_p), ctypes.cast(ctypes.byref(txin), _c_double_p))
    return wekp, txis.value, txin.value


# ------------------------------------------------------------------ ocean mixed layer


def oml_params(o) -> CrOmlParams:
    """CrOmlParams from a dict with hm, T1_rel, dT_12, rho_cp, d2, d4, ycexp, nval, sval, fluxn, fluxs."""
    P = CrOmlParams()
    for k in ("hm", "T1_rel", "dT_12", "rho_cp", "d2", "d4", "ycexp", "nval", "sval"):
        setattr(P, k, float(o[k]))
    P.fluxn, P.fluxs = int(bool(o["fluxn"])), int(bool(o["fluxs"]))
    return P


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def del2_T(cfg, t, north_flux=False, nval=0.0, south_flux=False, sval=0.0):
    P = make_params(cfg)
    t = _c(t)
    out = np.zeros_like(t)
    lib().cr_del2_T(ctypes.byref(P), _ptr(t), int(north_flux), float(nval), int(south_flux), float(sval), _ptr(out))
    return out


def ekman_uvw(cfg, hm, taux, tauy):
    """(uek (nyt,nxp), vek (nyp,nxt), wekt (nyt,nxt)) from the kinematic stress at p points."""
    P = make_params(cfg)
    nxp, nyp = P.nxp, P.nyp
    uek, vek, wekt = np.zeros((nyp - 1, nxp)), np.zeros((nyp, nxp - 1)), np.zeros((nyp - 1, nxp - 1))
    lib().cr_ekman_uvw(ctypes.byref(P), float(hm), _ptr(_c(taux)), _ptr(_c(tauy)), _ptr(uek), _ptr(vek), _ptr(wekt))
    return uek, vek, wekt


def adv_diff(cfg, o, p, uek, vek, data, datam):
    P, O = make_params(cfg), oml_params(o)
    out = np.zeros((P.nyp - 1, P.nxp - 1))
    lib().cr_adv_diff(ctypes.byref(P), ctypes.byref(O), _ptr(_c(p)), _ptr(_c(uek)), _ptr(_c(vek)), _ptr(_c(data)),
                      _ptr(_c(datam)), _ptr(out))
    return out


def map_T_to_P(cfg, t):
    P = make_params(cfg)
    out = np.zeros((P.nyp, P.nxp))
    lib().cr_map_T_to_P(ctypes.byref(P), _ptr(_c(t)), _ptr(out))
    return out


def oml_step(cfg, o, p1, taux, tauy, fnet, tdto, sst, sstm, compute_ent=True):
    """step_oml: returns (sst, sstm, ent1) after one step (inputs are not modified)."""
    P, O = make_params(cfg), oml_params(o)
    sst, sstm = _c(sst).copy(), _c(sstm).copy()
    ent1 = np.zeros((P.nyp, P.nxp))
    lib().cr_oml_step(ctypes.byref(P), ctypes.byref(O), _ptr(_c(p1)), _ptr(_c(taux)), _ptr(_c(tauy)), _ptr(_c(fnet)),
                      float(tdto), int(compute_ent), _ptr(sst), _ptr(sstm), _ptr(ent1))
    return sst, sstm, ent1


# ------------------------------------------------------------------ coupling


def cpl(cfg_o, cfg_a, xo0=0.0, yo0=0.0, xa0=0.0, ya0=0.0, hmo=100.0, hma=1000.0, rho_o=1000.0, rho_a=1.0,
        cdat=1.3e-3, tau_udiff=False) -> CrCpl:
    """The coupling description: load_grid (grid.F90:37-58) plus origins, depths, densities, drag."""
    g = CrCpl()
    lib().cr_load_grid(ctypes.byref(make_params(cfg_o)), ctypes.byref(make_params(cfg_a)), ctypes.byref(g))
    g.xo0, g.yo0, g.xa0, g.ya0, g.hmo, g.hma = xo0, yo0, xa0, ya0, hmo, hma
    g.rho_o, g.rho_a, g.cdat, g.tau_udiff = rho_o, rho_a, cdat, int(bool(tau_udiff))
    return g


def aml_params(a) -> CrAmlParams:
    A = CrAmlParams()
    for n in _AML_D:
        setattr(A, n, float(a[n]))
    for k, v in enumerate(a["aface"]):
        A.aface[k] = float(v)
    A.bface, A.dface = float(a["bface"]), float(a["dface"])
    return A


def bcuini(ndxr):
    """st((ndxr+1)^2, 16) as a numpy array [c][r] (column-major storage)."""
    st = np.zeros((16, (ndxr + 1) ** 2))
    lib().cr_bcuini(int(ndxr), _ptr(st))
    return st


def deriv_2bc(name, cfg, alpha, p):
    """dPdx_2bc, dPdy_2bc, dP2dx2_bc or dP2dy2_bc (numerics.F90:427-513) of a (nyp, nxp) field."""
    P = make_params(cfg)
    out = np.zeros((P.nyp, P.nxp))
    getattr(lib(), "cr_" + name)(ctypes.byref(P), float(alpha), _ptr(_c(p)), _ptr(out))
    return out


def new_bicubic(g, nxpo, nypo, st, u, ux, uy, uxy):
    out = np.zeros((nypo, nxpo))
    lib().cr_new_bicubic(ctypes.byref(g), int(nxpo), int(nypo), _ptr(_c(st)), _ptr(_c(u)), _ptr(_c(ux)), _ptr(_c(uy)),
                         _ptr(_c(uxy)), _ptr(out))
    return out


def xforc(cfg_o, cfg_a, g, pmo1, pma1):
    """(tauxo, tauyo) on the ocean p grid and (tauxa, tauya) on the atmosphere's (windstress.F90:102-316)."""
    Po, Pa = make_params(cfg_o), make_params(cfg_a)
    st = bcuini(g.ndxr)
    out = [np.zeros((Po.nyp, Po.nxp)) for _ in range(2)] + [np.zeros((Pa.nyp, Pa.nxp)) for _ in range(2)]
    lib().cr_xforc(ctypes.byref(Po), ctypes.byref(Pa), ctypes.byref(g), _ptr(st), _ptr(_c(pmo1)), _ptr(_c(pma1)),
                   *[_ptr(o) for o in out])
    return tuple(out)


def bilint(cfg_a, cfg_o, g, atmos):
    Pa, Po = make_params(cfg_a), make_params(cfg_o)
    out = np.zeros((Po.nyp - 1, Po.nxp - 1))
    lib().cr_bilint(ctypes.byref(Pa), ctypes.byref(Po), ctypes.byref(g), _ptr(_c(atmos)), _ptr(out))
    return out


def compute_forcing(cfg_o, cfg_a, g, a, sstm, astm, hmixam, eta1, dtop):
    """(fnetoc (ocean T), fnetat (atmos T)) of forcing.F90:18-99."""
    Po, Pa = make_params(cfg_o), make_params(cfg_a)
    fo = np.zeros((Po.nyp - 1, Po.nxp - 1))
    fa = np.zeros((Pa.nyp - 1, Pa.nxp - 1))
    lib().cr_compute_forcing(ctypes.byref(Po), ctypes.byref(Pa), ctypes.byref(g), ctypes.byref(aml_params(a)),
                             _ptr(_c(sstm)), _ptr(_c(astm)), _ptr(_c(hmixam)), _ptr(_c(eta1)), _ptr(_c(dtop)),
                             _ptr(fo), _ptr(fa))
    return fo, fa


def aml_step(cfg_a, g, a, p1, taux, tauy, eta, dtop, fnetat, xc1ast, tdta, ast, astm, hmixa, hmixam,
             compute_ent=True):
    """step_aml (amlsubs.F90:184-326): returns (ast, astm, hmixa, hmixam, ent1); inputs are not modified."""
    Pa = make_params(cfg_a)
    ast, astm, hmixa, hmixam = (_c(x).copy() for x in (ast, astm, hmixa, hmixam))
    ent1 = np.zeros((Pa.nyp, Pa.nxp))
    lib().cr_aml_step(ctypes.byref(Pa), ctypes.byref(g), ctypes.byref(aml_params(a)), _ptr(_c(p1)), _ptr(_c(taux)),
                      _ptr(_c(tauy)), _ptr(_c(eta)), _ptr(_c(dtop)), _ptr(_c(fnetat)), _ptr(_c(xc1ast)), float(tdta),
                      int(compute_ent), _ptr(ast), _ptr(astm), _ptr(hmixa), _ptr(hmixam), _ptr(ent1))
    return ast, astm, hmixa, hmixam, ent1
