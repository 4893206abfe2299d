"""An OpenQG experiment from its staged configuration files, run on the GPU.

`Experiment(stage_dir)` reads the files the reference's model reads from its
input directory (src/openqg.F90:60-218): glam.cdl (the mesh, src/glam/mesh.F90),
ocn_basin.cdl and atm_basin.cdl (src/basin.F90), windstress.cdl
(src/windstress/windstress.F90:40-75) and clock.cdl (src/types/clock.F90), as
CDL text (openqg_amd.cdl) rather than through ncgen.  It builds the basins
(init_box_from_mesh, box.F90:50-90; init_qg), the mixed layers, the radiation
and the coupler, sets the initial state ("zero", "rbal" or a restart file,
init_dynamic_state, openqg.F90:221-330) and `run(outdir)` runs the time loop of
openqg.F90:453-512 (openqg_amd.driver.run) writing the restart and lastday files.

Supported: the coupled double gyre (both basins with their mixed layers and the
coupled stress) and ocean-only basins with a fixed stress read from a netCDF
file (tauxo, tauyo; load_tau_from_file, ekman.F90:85-102).  Topography other
than "flat" and the netCDF diagnostics (monitors, averages, subsampling) are
not part of this harness.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from .cdl import read_cdl
from .coupled import Coupler, load_grid
from .driver import Basin, Clock, Valids, run
from .mixed import OceanMixedLayer, init_temp_from_rad, zonal_temp
from .qg import Box, init_qg
from .radiation import entrain_coeff, load_rad, radiat, xc1ast
from .state import restart_aml, restart_oml, restart_qg

PI = 3.14159265358979324
R0 = 6378.1e3                                        # constants.F90:22 (a single-precision literal there)
OMEGA = 2 * PI / ((23 * 60 + 56) * 60 + 4.098903691)  # constants.F90:23-24: one sidereal day


@dataclass
class Mesh:
    """mesh_type (mesh.F90:12-32) with fnot, beta from the latitude (compute_derived_values)."""

    nxt: int
    nyt: int
    dx: float
    dy: float
    x0: float
    y0: float
    cyclic: bool
    x1: float
    lat: float

    def __post_init__(self):
        phi = PI * self.lat / 180.0
        if self.cyclic:
            R = (self.nxt * self.dx) / (2 * PI * math.cos(phi))
        elif self.x1 == 0.0:
            R = R0 * math.cos(phi)
        else:
            R = self.x1 / (2 * PI * math.cos(phi))
        self.fnot = 2 * OMEGA * math.sin(phi)
        self.beta = 2 * OMEGA * math.cos(phi) / R


def box_from_mesh(mesh: Mesh, cfg: dict, dz_sign: int) -> Box:
    """init_box_from_mesh (box.F90:50-90)."""
    nx, ny, ndx, ndy = int(cfg["nx"]), int(cfg["ny"]), int(cfg["ndx"]), int(cfg["ndy"])
    h = cfg["h"] if isinstance(cfg["h"], list) else [cfg["h"]]
    return Box(nx * ndx + 1, ny * ndy + 1, len(h), mesh.dx / ndx, mesh.dy / ndy, mesh.fnot, mesh.beta, list(h),
               cyclic=(nx == mesh.nxt), dz_sign=dz_sign, hm=float(cfg["hm"]), x0=(int(cfg["nx1"]) - 1) * mesh.dx,
               y0=(int(cfg["ny1"]) - 1) * mesh.dy)


def _lst(v, n):
    return [float(x) for x in (v if isinstance(v, list) else [v])][:n]


def _valids(cfg: dict, clk: Clock):
    if not int(cfg.get("check_valids", 0)):
        return None
    return Valids(clk.days_to_steps(float(cfg["dtvalid"])), float(cfg["max_p"]), float(cfg["max_q"]))


class Experiment:
    def __init__(self, stage_dir: str):
        self.stage = stage_dir
        path = lambda n: os.path.join(stage_dir, n)
        g = read_cdl(path("glam.cdl"))
        self.mesh = Mesh(int(g["nxt"]), int(g["nyt"]), float(g["dx"]), float(g["dy"]), float(g["x0"]), float(g["y0"]),
                         bool(int(g["cyclic"])), float(g.get("x1", 0.0)), float(g["lat"]))
        c = read_cdl(path("clock.cdl"))
        self.clk = Clock(float(c["dta"]), int(c["nstr"]), float(c["tini"]), float(c["trun"]), float(c["resday"]))
        self.wind = read_cdl(path("windstress.cdl"))
        use_ocn = int(g.get("use_ocean", 1)) and os.path.exists(path("ocn_basin.cdl"))
        use_atm = int(g.get("use_atmos", 1)) and os.path.exists(path("atm_basin.cdl"))
        self.ocfg = read_cdl(path("ocn_basin.cdl")) if use_ocn else None
        self.acfg = read_cdl(path("atm_basin.cdl")) if use_atm else None
        self.ocn = self.atm = self.oml = self.cpl = None
        for cfg in (self.ocfg, self.acfg):
            if cfg is not None and str(cfg.get("topog", "flat")) != "flat":
                raise NotImplementedError(f"topography '{cfg['topog']}': only flat basins are supported")
        if self.ocfg is not None:
            self.go = box_from_mesh(self.mesh, self.ocfg, -1)
            self.ocn = self._qg(self.ocfg, self.go, tau_sign=1)
        if self.acfg is not None:
            self.ga = box_from_mesh(self.mesh, self.acfg, 1)
            self.atm = self._qg(self.acfg, self.ga, tau_sign=-1)
        self._mixed_and_coupling()

    def _qg(self, cfg, b, tau_sign):
        nl = b.nl
        return init_qg(b, float(cfg["bcco"]), float(cfg["delek"]), _lst(cfg["ah2"], nl), _lst(cfg["ah4"], nl),
                       _lst(cfg["gp"], nl - 1), float(cfg["rho"]), tau_sign)

    def _mixed_and_coupling(self):
        """init_mixed_and_qg (openqg.F90:190-218) and the stress (windstress.F90:77-100)."""
        o, a = self.ocfg, self.acfg
        coupled_wind = int(self.wind.get("coupled", 0)) == 1
        use_aml = a is not None and int(a.get("use_aml", 0))
        use_oml = o is not None and int(o.get("use_oml", 0))
        if use_aml or use_oml or coupled_wind:
            if not (use_aml and use_oml and coupled_wind):
                raise NotImplementedError("mixed layers and the coupled stress are supported together only")
            ga, go = self.ga, self.go
            zopt = [float(a["zopt_m"])] + _lst(a["zopt"], ga.nl)
            self.rad = rad = load_rad(ga, float(a["fsbar"]), float(a["fspamp"]), zopt, float(a["gamma"]),
                                      float(a["xlamda"]), _lst(a["T_abs"], ga.nl))
            rbtmat, rbtmoc, ocn_tmbar, atm_tmbar = radiat(ga, rad)
            self.attemp = attemp = init_temp_from_rad(ga, ga, rad.tabsat[0], rad.tabsat[1], atm_tmbar, rbtmat,
                                                      rad.fspco, float(a["rho"]), float(a["cp"]))
            if not int(o.get("use_rad_temp", 1)):
                raise NotImplementedError("use_rad_temp = 0 (openqg.F90:214-216)")
            self.octemp = octemp = init_temp_from_rad(go, ga, float(o["T1_abs"]), float(o["T2_abs"]), ocn_tmbar,
                                                      rbtmoc, rad.fspco, float(o["rho"]), float(o["cp"]))
            self.oml = OceanMixedLayer(self.ocn, octemp, bool(int(o["flux_n"])), bool(int(o["flux_s"])),
                                       float(o["st2d"]), float(o["st4d"]), float(o["ycexp"]), hm=go.hm)
            aface, bface, dface = entrain_coeff(ga.nl, attemp, rad)
            prm = dict(xo0=go.x0, yo0=go.y0, xa0=ga.x0, ya0=ga.y0, hmo=go.hm, hma=ga.hm, rho_o=float(o["rho"]),
                       rho_a=float(a["rho"]), cdat=float(self.wind["cdat"]),
                       tau_udiff=bool(int(self.wind.get("tau_udiff", 0))), D0up=rad.D0up, Dmdown=rad.Dmdown,
                       Dup0=float(rad.Dup[0]), xlamda=rad.xlamda, fspco=rad.fspco, Aup00=float(rad.Aup[0, 0]),
                       Adown10=float(rad.Adown[0, 0]), Adown11=float(rad.Adown[0, 1]), hmamin=float(a["hmamin"]),
                       hmadmp=float(a["hmadmp"]), xcexp=float(a["xcexp"]), T1_rel=attemp.T1_rel,
                       dT_12=attemp.dT_12, rho_cp=attemp.rho_cp, at2d=float(a["st2d"]), at4d=float(a["st4d"]),
                       ahmd=float(a["ahmd"]), aface=aface, bface=bface, dface=dface)
            self.cpl = Coupler(self.oml, self.atm, prm, xc1ast=xc1ast(ga, attemp))
            self.grid = load_grid(go, ga)
        elif self.ocn is not None:
            fn = str(self.wind.get("ocn_stress", ""))
            if not fn:
                raise ValueError("an uncoupled ocean needs windstress.cdl's ocn_stress file")
            from scipy.io import netcdf_file

            with netcdf_file(os.path.join(self.stage, fn), "r", mmap=False) as f:
                tx = np.array(f.variables["tauxo"][:], dtype=np.float64)
                ty = np.array(f.variables["tauyo"][:], dtype=np.float64)
            self.ocn.set_wind(tx, ty)

    def _field_rows(self, box, temp):
        rows = np.array([zonal_temp(box, temp, j) for j in range(1, box.nyt + 1)])
        return rows[:, None] + np.zeros((1, box.nxt))

    def init_state(self):
        """init_dynamic_state (openqg.F90:221-330)."""
        for cfg, qg, kind in ((self.ocfg, self.ocn, "ocn"), (self.acfg, self.atm, "atm")):
            if qg is None:
                continue
            st = str(cfg.get("qg_state", "zero"))
            if st == "zero":
                z = np.zeros((qg.b.nl, qg.b.nyp, qg.b.nxp))
                qg.set_state(p=z, pm=z)
                qg.init_state()
            elif st == "rbal" and kind == "atm":
                b, rad = qg.b, self.rad
                plfac = [0.0]
                for k in range(2, b.nl + 1):
                    plfac.append(plfac[-1] - qg.gp[k - 2] * rad.rbetat[k - 2])
                fs = rad.fspco * 0.5 * np.sin(PI * b.yprel / b.yl)   # fsprim (radsubs.F90:97-114)
                p = np.array(plfac)[:, None, None] * fs[None, :, None] + np.zeros((1, 1, b.nxp))
                qg.set_state(p=p, pm=p)
                qg.init_state()
            else:
                restart_qg(self.stage, st, qg)
        if self.oml is not None:
            st = str(self.ocfg.get("oml_state", "zero"))
            if st == "zero":
                z = np.zeros(self.oml.shape)
                self.oml.set_state(z, z)
            elif st == "rbal":
                t = self._field_rows(self.go, self.octemp)
                self.oml.set_state(t, t)
            else:
                restart_oml(self.stage, st, self.oml)
        if self.cpl is not None:
            st = str(self.acfg.get("aml_state", "zero"))
            shp = self.cpl.tshape_a
            if st == "zero":
                z, h = np.zeros(shp), np.full(shp, self.ga.hm)
                self.cpl.set_aml_state(z, z, h, h)
            elif st == "rbal":
                t, h = self._field_rows(self.ga, self.attemp), np.full(shp, self.ga.hm)
                self.cpl.set_aml_state(t, t, h, h)
            else:
                restart_aml(self.stage, st, self.cpl)

    def run(self, outdir: str, monitor=None) -> int:
        """The time loop (openqg.F90:453-512) from the initial state; returns the last step."""
        os.makedirs(outdir, exist_ok=True)
        self.init_state()
        if self.cpl is not None:   # the first exchange (openqg.F90:456-478)
            self.cpl.xforc()
            self.cpl.forcing()
        ocn = Basin(str(self.ocfg["name"]), self.ocn, _valids(self.ocfg, self.clk)) if self.ocn is not None else None
        atm = Basin(str(self.acfg["name"]), self.atm, _valids(self.acfg, self.clk)) if self.atm is not None else None
        return run(self.clk, outdir, ocn=ocn, atm=atm, monitor=monitor, coupler=self.cpl)

    def close(self):
        for q in (self.ocn, self.atm):
            if q is not None:
                q.close()
