"""Host mirror of OpenQG's QG component over the HIP library.

`Box`, `Topog` and `QG` carry the fields of box_type (src/glam/box.F90:11-41),
topog_type (src/glam/topog.F90:16-28) and qg_type (src/qg/qg.F90:17-53) that
the time step reads; `init_qg` mirrors qg.F90:60.  The dynamic state lives on
the GPU; the `q`, `qm`, `p`, `pm` properties copy it back as numpy arrays of
shape (nl, nyp, nxp) (the C view of the Fortran (nxp, nyp, nl) arrays).
Fields may also be passed as torch tensors (CPU or GPU, float64, contiguous).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import OQG_DEVICE, OQG_HOST, OqgDesc, check, load


@dataclass
class Box:
    """box_type: p-grid size, spacing, layers and Coriolis parameters."""

    nxp: int
    nyp: int
    nl: int
    dx: float
    dy: float
    fnot: float
    beta: float
    h: Sequence[float]
    cyclic: bool = False
    dz_sign: int = -1
    hm: float = 0.0
    x0: float = 0.0   # position of p point (1,1) in its mesh (box.F90:62-63)
    y0: float = 0.0

    @property
    def nxt(self):
        return self.nxp - 1

    @property
    def nyt(self):
        return self.nyp - 1

    @property
    def xl(self):
        return self.nxt * self.dx          # box.F90:135

    @property
    def yl(self):
        return self.nyt * self.dy          # box.F90:136

    @property
    def yprel(self):
        j = np.arange(self.nyp, dtype=np.float64)
        return -0.5 * self.yl + j * self.dy  # box.F90:144

    @property
    def xp(self):
        return self.x0 + np.arange(self.nxp, dtype=np.float64) * self.dx   # box.F90:154

    @property
    def yp(self):
        return self.y0 + np.arange(self.nyp, dtype=np.float64) * self.dy   # box.F90:143


@dataclass
class Topog:
    """topog_type: dynamic topography ddyn = f0/H(k_topo) * dtop and k_topo (1-based)."""

    ddyn: Optional[np.ndarray] = None
    k_topo: Optional[int] = None


def _as_f64_contig(a):
    """Return (object keeping memory alive, pointer, kind) for numpy or torch input."""
    if a is None:
        return None, None, OQG_HOST
    try:
        import torch  # noqa: F401

        if isinstance(a, torch.Tensor):
            if a.dtype != torch.float64:
                raise TypeError("fields must be float64")
            t = a.contiguous()
            if t.is_cuda:
                torch.cuda.current_stream(t.device).synchronize()
                return t, ctypes.c_void_p(t.data_ptr()), OQG_DEVICE
            return t, ctypes.c_void_p(t.data_ptr()), OQG_HOST
    except ImportError:
        pass
    arr = np.ascontiguousarray(a, dtype=np.float64)
    return arr, arr.ctypes.data_as(ctypes.c_void_p), OQG_HOST


class QG:
    """One basin resident on the GPU (qg_type + liboqg context)."""

    def __init__(self, b: Box, bcco: float, delek: float, ah2, ah4, gp, rho: float = 1000.0,
                 tau_sign: int = 1, topo: Optional[Topog] = None, flags: int = 0, precompute=None):
        lib = load()
        self.b = b
        self.bcco, self.delek, self.rho, self.tau_sign = float(bcco), float(delek), float(rho), int(tau_sign)
        self.ah2 = [float(x) for x in ah2]
        self.ah4 = [float(x) for x in ah4]
        self.gp = [float(x) for x in gp][: b.nl - 1]
        self.topo = topo or Topog()
        self.flags = int(flags)
        d = OqgDesc()
        d.nxp, d.nyp, d.nl, d.cyclic = b.nxp, b.nyp, b.nl, int(bool(b.cyclic))
        d.dx, d.dy, d.fnot, d.beta = b.dx, b.dy, b.fnot, b.beta
        for k in range(b.nl):
            d.h[k], d.ah2[k], d.ah4[k] = b.h[k], self.ah2[k], self.ah4[k]
        for k in range(b.nl - 1):
            d.gp[k] = self.gp[k]
        d.bcco, d.delek = self.bcco, self.delek
        d.dz_sign, d.tau_sign = int(b.dz_sign), self.tau_sign
        d.k_topo = int(self.topo.k_topo or 0)
        d.flags = self.flags
        keep = []

        def ptr(a):
            if a is None:
                return None
            arr = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        if self.topo.ddyn is not None:
            d.ddyn = ptr(self.topo.ddyn)
        if precompute:
            # Fortran column-major (nl,nl) arrays, as held in qg%mod / qg%hom
            for name in ("amat", "ctl2m", "ctm2l"):
                if name in precompute:
                    setattr(d, name, ptr(np.asarray(precompute[name]).T))
            for name in ("rdm2", "hom_sol_bc", "hom_bt", "hom_bc1", "hom_bc2", "hom_coef"):
                if name in precompute:
                    setattr(d, name, ptr(precompute[name]))
        h = ctypes.c_void_p()
        rc = lib.oqg_create(ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            raise _lib.OqgError(f"oqg_create failed (code {rc}): {lib.oqg_last_error(None).decode()}")
        self._h = h
        self._lib = lib

    # ------------------------------------------------------------ lifecycle
    def _attach(self, obj):
        """Register an object bound to this basin's context (a mixed layer); it is
        closed before the context is destroyed."""
        import weakref

        self.__dict__.setdefault("_bound", []).append(weakref.ref(obj))

    def close(self):
        for r in reversed(self.__dict__.pop("_bound", [])):   # the last bound first (it may use earlier ones)
            o = r()
            if o is not None:
                o.close()
        if getattr(self, "_h", None) is not None:
            self._lib.oqg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.b.nl, self.b.nyp, self.b.nxp)

    def to_dict(self) -> dict:
        """Parameters as a plain dict (the form the test oracle takes)."""
        b = self.b
        return dict(nxp=b.nxp, nyp=b.nyp, nl=b.nl, cyclic=bool(b.cyclic), dx=b.dx, dy=b.dy, fnot=b.fnot,
                    beta=b.beta, h=list(b.h), gp=list(self.gp), ah2=list(self.ah2), ah4=list(self.ah4),
                    bcco=self.bcco, delek=self.delek, dz_sign=b.dz_sign, tau_sign=self.tau_sign,
                    k_topo=self.topo.k_topo, ddyn=self.topo.ddyn)

    # ------------------------------------------------------------ state
    def set_state(self, q=None, qm=None, p=None, pm=None):
        objs = [_as_f64_contig(a) for a in (q, qm, p, pm)]
        kinds = {o[2] for o in objs if o[0] is not None}
        if len(kinds) > 1:
            raise ValueError("pass all fields on the host or all on the device")
        kind = kinds.pop() if kinds else OQG_HOST
        for o in objs:
            if o[0] is not None and tuple(o[0].shape) != self.shape:
                raise ValueError(f"field shape {tuple(o[0].shape)} != {self.shape}")
        check(self._lib.oqg_set_state(self._h, *(o[1] for o in objs), kind), self._h, "oqg_set_state")

    def get_state(self):
        outs = [np.empty(self.shape) for _ in range(4)]
        ptrs = [o.ctypes.data_as(ctypes.c_void_p) for o in outs]
        check(self._lib.oqg_get_state(self._h, *ptrs, OQG_HOST), self._h, "oqg_get_state")
        return tuple(outs)

    def get_state_torch(self, device="cuda"):
        """Copy q, qm, p, pm into new torch tensors on `device` (device-to-device)."""
        import torch

        outs = [torch.empty(self.shape, dtype=torch.float64, device=device) for _ in range(4)]
        torch.cuda.current_stream(outs[0].device).synchronize()
        ptrs = [ctypes.c_void_p(t.data_ptr()) for t in outs]
        check(self._lib.oqg_get_state(self._h, *ptrs, OQG_DEVICE), self._h, "oqg_get_state")
        return tuple(outs)

    def _field(self, which):
        out = np.empty(self.shape)
        for k in range(self.b.nl):
            check(self._lib.oqg_get_layer(self._h, which, k + 1,
                                          out[k].ctypes.data_as(ctypes.c_void_p), OQG_HOST),
                  self._h, "oqg_get_layer")
        return out

    q = property(lambda self: self._field(_lib.FIELD_Q))
    qm = property(lambda self: self._field(_lib.FIELD_QM))
    p = property(lambda self: self._field(_lib.FIELD_P))
    pm = property(lambda self: self._field(_lib.FIELD_PM))

    def get_layer(self, which: int, k: int) -> np.ndarray:
        out = np.empty((self.b.nyp, self.b.nxp))
        check(self._lib.oqg_get_layer(self._h, which, k, out.ctypes.data_as(ctypes.c_void_p), OQG_HOST),
              self._h, "oqg_get_layer")
        return out

    def get_constr(self) -> np.ndarray:
        n = self._lib.oqg_constr_len(self._h)
        out = np.zeros(n)
        check(self._lib.oqg_get_constr(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              self._h, "oqg_get_constr")
        return out

    def set_constr(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(self._lib.oqg_set_constr(self._h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              self._h, "oqg_set_constr")

    def init_state(self):
        """init_pv + init_constraint + init_mass_constr (openqg.F90:239-243)."""
        check(self._lib.oqg_init_state(self._h), self._h, "oqg_init_state")

    # ------------------------------------------------------------ stepping
    def set_forcing(self, ent, wekp, txis: float = 0.0, txin: float = 0.0):
        ent_o = _as_f64_contig(ent)
        wek_o = _as_f64_contig(wekp)
        n_ent = 1 if ent_o[0] is None or ent_o[0].ndim == 2 else int(ent_o[0].shape[0])
        if ent_o[0] is not None and wek_o[0] is not None and ent_o[2] != wek_o[2]:
            raise ValueError("ent and wekp must both be on the host or both on the device")
        kind = ent_o[2] if ent_o[0] is not None else wek_o[2]
        check(self._lib.oqg_set_forcing(self._h, ent_o[1], n_ent, wek_o[1], float(txis), float(txin), kind),
              self._h, "oqg_set_forcing")

    def set_wind(self, taux, tauy):
        """ekman_from_tau on the device (ekman.F90:53-83): the resident wekp (and a
        channel's txis/txin) from the kinematic stress (nyp, nxp) at p points."""
        tx_o = _as_f64_contig(taux)
        ty_o = _as_f64_contig(tauy)
        if tx_o[2] != ty_o[2]:
            raise ValueError("taux and tauy must both be on the host or both on the device")
        for o in (tx_o, ty_o):
            if tuple(o[0].shape) != (self.b.nyp, self.b.nxp):
                raise ValueError(f"stress shape {tuple(o[0].shape)} != {(self.b.nyp, self.b.nxp)}")
        check(self._lib.oqg_set_wind(self._h, tx_o[1], ty_o[1], tx_o[2]), self._h, "oqg_set_wind")

    def diagnose(self, dt: float) -> dict:
        """qg_diagno's state energetics and averages, scan_cfl's maxima and the
        count of non-finite values, computed on the device (oqg_diagnose); in a
        cyclic basin also check_continuity's ermas and emfr (tdt = 2 dt)."""
        d = _lib.OqgDiag()
        check(self._lib.oqg_diagnose(self._h, float(dt), self.rho, ctypes.byref(d)), self._h, "oqg_diagnose")
        nl = self.b.nl
        out = {}
        for name, _ in _lib.OqgDiag._fields_:
            v = getattr(d, name)
            if isinstance(v, float):
                out[name] = v
            else:
                out[name] = np.array(v[: nl - 1] if name in ("etam", "et2m") else v[:nl])
        if self.b.cyclic:
            out["ermas"], out["emfr"] = self._continuity(out["pint"], 2.0 * float(dt))
        return out

    def invert(self):
        """p from the current q (oqg_invert): modes, Helmholtz solves and the
        mass correction with the current dpi; q's boundary PV by ocqbdy; then
        qm = q and pm = p."""
        check(self._lib.oqg_invert(self._h), self._h, "oqg_invert")

    def diagnose_energy(self) -> dict:
        """The rest of qg_diagno's energetics on the device
        (oqg_diagnose_energy): ah2d, ah4d, the jet position and value per
        layer; btdg, utau, wepm, wapm, entm, enam, pken."""
        e = _lib.OqgEnergy()
        check(self._lib.oqg_diagnose_energy(self._h, self.rho, ctypes.byref(e)), self._h, "oqg_diagnose_energy")
        nl = self.b.nl
        out = {}
        for name, _ in _lib.OqgEnergy._fields_:
            v = getattr(e, name)
            out[name] = v if isinstance(v, float) else np.array(v[:nl])
        out["pos"] = out["pos"].astype(int)
        return out

    def _continuity(self, int_p, tdt):
        """check_continuity (qg_monitor.F90:571-604, cyclic basins only): the
        interface volume from the layer integrals int_P_dA(p_k) against the
        constraint's dpi.  Returns (ermas, emfr) per interface."""
        b = self.b
        nl = b.nl
        area = b.xl * b.yl
        int_p = np.asarray(int_p)
        dpi = self.get_constr()[: nl - 1]
        ermas, emfr = np.zeros(nl - 1), np.zeros(nl - 1)
        for k in range(nl - 1):
            est1 = b.dz_sign * (int_p[k] - int_p[k + 1])
            est2 = dpi[k]
            esum = abs(est1) + abs(est2)
            ermas[k] = est1 - est2
            emfr[k] = 2.0 * ermas[k] / esum if esum > 1.0e-13 * area * tdt * self.gp[k] else 0.0
        return ermas, emfr

    def get_forcing(self):
        """(wekp (nyp, nxp), txis, txin) as resident on the device."""
        w = np.empty((self.b.nyp, self.b.nxp))
        txi = (ctypes.c_double * 2)()
        check(self._lib.oqg_get_forcing(self._h, w.ctypes.data_as(ctypes.c_void_p), txi, OQG_HOST),
              self._h, "oqg_get_forcing")
        return w, float(txi[0]), float(txi[1])

    def get_wind(self):
        """The resident kinematic wind stress (taux, tauy), each (nyp, nxp)."""
        tx, ty = np.empty((self.b.nyp, self.b.nxp)), np.empty((self.b.nyp, self.b.nxp))
        check(self._lib.oqg_get_wind(self._h, tx.ctypes.data_as(ctypes.c_void_p), ty.ctypes.data_as(ctypes.c_void_p),
                                     OQG_HOST), self._h, "oqg_get_wind")
        return tx, ty

    def get_ent(self):
        """The resident layer 1/2 entrainment (nyp, nxp)."""
        e = np.empty((self.b.nyp, self.b.nxp))
        check(self._lib.oqg_get_ent(self._h, e.ctypes.data_as(ctypes.c_void_p), OQG_HOST), self._h, "oqg_get_ent")
        return e

    def step(self, tdt: float, ent=None, wekp=None, txis: float = 0.0, txin: float = 0.0):
        """step_qg (core.F90:16).  With ent and wekp None the resident forcing is used."""
        if ent is None and wekp is None:
            check(self._lib.oqg_step(self._h, float(tdt), None, 0, None, 0.0, 0.0), self._h, "oqg_step")
            return
        self.set_forcing(ent, wekp, txis, txin)
        check(self._lib.oqg_step(self._h, float(tdt), None, 0, None, 0.0, 0.0), self._h, "oqg_step")

    def suppress(self):
        """suppress_comp_mode (driver.F90:174-189)."""
        check(self._lib.oqg_suppress(self._h), self._h, "oqg_suppress")

    def run(self, tdt: float, nsteps: int, step0: int = 0, supp_every: int = 25):
        """Ocean-only time loop on the resident forcing (openqg.F90:483-509)."""
        check(self._lib.oqg_run(self._h, float(tdt), int(nsteps), int(step0), int(supp_every)), self._h, "oqg_run")

    def plan_run(self, tdt: float, nsteps: int, step0: int = 0, supp_every: int = 25):
        """Capture, without running, the graphs that `run` with the same arguments
        replays from the current time level (its short head and tail included)."""
        check(self._lib.oqg_plan_run(self._h, float(tdt), int(nsteps), int(step0), int(supp_every)), self._h,
              "oqg_plan_run")

    def sync(self):
        check(self._lib.oqg_sync(self._h), self._h, "oqg_sync")

    def solve_inhomog(self, m: int, rhs) -> np.ndarray:
        """solve_inhomog_eqn (inhomog.F90:156) for mode m (1-based)."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.empty_like(rhs)
        check(self._lib.oqg_solve_inhomog(self._h, int(m), rhs.ctypes.data_as(ctypes.c_void_p),
                                          out.ctypes.data_as(ctypes.c_void_p), OQG_HOST),
              self._h, "oqg_solve_inhomog")
        return out

    def modes(self):
        """(amat, ctl2m, ctm2l, rdm2) with matrices indexed [row][col] as in Fortran (i,j)."""
        nl = self.b.nl
        mats = [np.zeros(nl * nl) for _ in range(3)]
        rdm2 = np.zeros(nl)
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(self._lib.oqg_get_modes(self._h, P(mats[0]), P(mats[1]), P(mats[2]), P(rdm2)), self._h, "oqg_get_modes")
        return tuple(m.reshape(nl, nl).T.copy() for m in mats) + (rdm2,)

    def kernel_times(self, tdt: float, reps: int = 20):
        """Advance one step and return each kernel's average duration (ms) over
        `reps` back-to-back launches between two HIP events on the library stream."""
        ms = (ctypes.c_double * 8)()
        names = ctypes.create_string_buffer(256)
        n = self._lib.oqg_kernel_times(self._h, float(tdt), int(reps), ms, 8, names, 256)
        if n < 0:
            check(-n, self._h, "oqg_kernel_times")
        keys = names.value.decode().split("\n")[:n]
        return {k: ms[i] for i, k in enumerate(keys)}

    @property
    def tile_rows(self) -> int:
        """Rows of the vorticity kernel's 64-column tiles (4, 8 or 16)."""
        return int(self._lib.oqg_get_tile_rows(self._h))

    @tile_rows.setter
    def tile_rows(self, rows: int):
        check(self._lib.oqg_set_tile_rows(self._h, int(rows)), self._h, "oqg_set_tile_rows")

    def tune(self, tdt: float, trial_steps: int = 50, supp_every: int = 25):
        """Time `trial_steps` steps with each tile height (the state is
        restored), keep the fastest; returns (rows, {rows: ms})."""
        ms = (ctypes.c_double * 3)()
        best = self._lib.oqg_tune(self._h, float(tdt), int(trial_steps), int(supp_every), ms)
        if best < 0:
            check(-best, self._h, "oqg_tune")
        return int(best), dict(zip((4, 8, 16), list(ms)))

    def set_stream(self, stream_handle: int | None):
        check(self._lib.oqg_set_stream(self._h, ctypes.c_void_p(stream_handle) if stream_handle else None),
              self._h, "oqg_set_stream")


def init_qg(go: Box, bcco, delek, ah2, ah4, gp, rho, tau_sign, topo: Optional[Topog] = None, **kw) -> QG:
    """init_qg (src/qg/qg.F90:60-87): a resting basin with derived modes and Helmholtz tables."""
    return QG(go, bcco, delek, ah2, ah4, gp, rho, tau_sign, topo, **kw)
