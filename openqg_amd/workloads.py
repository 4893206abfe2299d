"""Synthetic workloads: the double-gyre ocean boxes and the atmosphere channel.

Each workload gives init_qg arguments, a time step, a run length, static wind
forcing and a seeded initial (p, pm).  q and qm are then derived on the device
by init_pv (vorsubs.F90:17), exactly as a restart from p would be
(openqg.F90:229-243).  All inputs are numpy float64.

Physics (ocean): 3 layers H = 300/1100/2600 m, g' = 0.05/0.025 m/s^2,
f0 = 1e-4 1/s, beta = 2e-11 1/(m s), bcco = 0.5, delek = 2 m, flat bottom,
tau_x = tau0 (1 - cos(2 pi y/L)) / rho0 with rho0 = 1000 kg/m^3.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .forcing import double_gyre_tau, ekman_wekp

RHO0 = 1000.0
F0 = 1.0e-4
BETA = 2.0e-11
OCEAN_LAYERS = {
    2: dict(h=[1000.0, 3000.0], gp=[0.02]),
    3: dict(h=[300.0, 1100.0, 2600.0], gp=[0.05, 0.025]),
    4: dict(h=[300.0, 700.0, 1200.0, 1800.0], gp=[0.05, 0.03, 0.02]),
}


@dataclass
class Workload:
    name: str
    nxp: int
    nyp: int
    nl: int
    dx: float
    dt: float
    days: float
    cyclic: bool = False
    h: list = field(default_factory=list)
    gp: list = field(default_factory=list)
    ah2: list = field(default_factory=list)
    ah4: list = field(default_factory=list)
    bcco: float = 0.5
    delek: float = 2.0
    dz_sign: int = -1
    tau_sign: int = 1
    tau0: float = 0.1
    fnot: float = F0
    beta: float = BETA
    init: str = "white"       # "white": iid PV noise; "k3": k^-3 spectrum field; "splitmix": q from SplitMix64
    q_rms: float = 1.0e-7
    seed: int = 1
    supp_every: int = 25

    @property
    def nsteps(self) -> int:
        return int(round(self.days * 86400.0 / self.dt))

    @property
    def tdt(self) -> float:
        return 2.0 * self.dt

    def params(self) -> dict:
        """Parameter dict (the oracle's input form)."""
        return dict(nxp=self.nxp, nyp=self.nyp, nl=self.nl, cyclic=self.cyclic, dx=self.dx, dy=self.dx,
                    fnot=self.fnot, beta=self.beta, h=list(self.h), gp=list(self.gp), ah2=list(self.ah2),
                    ah4=list(self.ah4), bcco=self.bcco, delek=self.delek, dz_sign=self.dz_sign,
                    tau_sign=self.tau_sign)

    def make_qg(self, **kw):
        from .qg import Box, init_qg

        b = Box(self.nxp, self.nyp, self.nl, self.dx, self.dx, self.fnot, self.beta, list(self.h),
                cyclic=self.cyclic, dz_sign=self.dz_sign)
        return init_qg(b, self.bcco, self.delek, self.ah2, self.ah4, self.gp, RHO0, self.tau_sign, **kw)

    def forcing(self):
        """(ent (nl-1, nyp, nxp), wekp (nyp, nxp), txis, txin); ent is zero for ocean-only runs."""
        tx, ty = double_gyre_tau(self.nyp, self.nxp, self.dx, self.tau0, RHO0)
        wekp, txis, txin = ekman_wekp(self.nxp, self.nyp, self.dx, self.dx, self.fnot, self.cyclic, tx, ty)
        ent = np.zeros((max(self.nl - 1, 1), self.nyp, self.nxp))
        return ent, wekp, txis, txin

    def initial_p(self):
        """Seeded p (= pm) whose PV anomaly has rms ~q_rms per layer."""
        if self.init == "splitmix":
            raise ValueError(f"{self.name} starts from q (Workload.start), not from a seeded p")
        return initial_p(self)

    def start(self, qg):
        """Put the initial state on the basin `qg`.  From a seeded p: p = pm,
        then init_pv / init_constraint / init_mass_constr (QG.init_state).
        init = "splitmix": q = qm = beta yprel + dq (dq = q_rms N(0, 1) from
        normal_deviates(seed) on the interior points, x fastest, then rows, then
        layers), p = pm = 0, dpi = 0, then p from one inversion (QG.invert)."""
        if self.init == "splitmix":
            z = np.zeros((self.nl, self.nyp, self.nxp))
            q = initial_q(self)
            qg.set_state(q=q, qm=q, p=z, pm=z)
            qg.set_constr(np.zeros(len(qg.get_constr())))
            qg.invert()
        else:
            p = self.initial_p()
            qg.set_state(p=p, pm=p)
            qg.init_state()


# SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
# generators", OOPSLA 2014; the finaliser constants of Stafford's Mix13).
SM_GAMMA = 0x9E3779B97F4A7C15
SM_M1 = 0xBF58476D1CE4E5B9
SM_M2 = 0x94D049BB133111EB


def splitmix64(seed: int, n: int) -> np.ndarray:
    """The first n outputs (uint64) of SplitMix64 started from state `seed`:
    state += gamma; z = state; z = (z ^ z >> 30) m1; z = (z ^ z >> 27) m2;
    output z ^ z >> 31, all modulo 2^64."""
    u = np.uint64
    with np.errstate(over="ignore"):
        s = u(seed % 2 ** 64) + np.arange(1, n + 1, dtype=np.uint64) * u(SM_GAMMA)
        z = (s ^ (s >> u(30))) * u(SM_M1)
        z = (z ^ (z >> u(27))) * u(SM_M2)
    return z ^ (z >> u(31))


def normal_deviates(seed: int, n: int) -> np.ndarray:
    """n standard normal deviates by the Box-Muller transform of SplitMix64
    uniforms u = (z >> 11) 2^-53 in [0, 1), taken in pairs (u1, u2):
    r = sqrt(-2 ln(1 - u1)), then r cos(2 pi u2) and r sin(2 pi u2)."""
    m = (n + 1) // 2
    z = splitmix64(seed, 2 * m)
    uni = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u1, u2 = uni[0::2], uni[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def initial_q(w: Workload) -> np.ndarray:
    """q = beta yprel + dq, dq = q_rms N(0, 1) on the interior points (x
    fastest, then rows, then layers), beta yprel on the boundary ring (which
    the inversion's ocqbdy replaces).  yprel = -yl/2 + j dy (box.F90:144)."""
    nl, nyp, nxp = w.nl, w.nyp, w.nxp
    yprel = -0.5 * (nyp - 1) * w.dx + np.arange(nyp) * w.dx
    q = np.broadcast_to((w.beta * yprel)[None, :, None], (nl, nyp, nxp)).copy()
    nyi, nxi = nyp - 2, nxp - 2
    dq = w.q_rms * normal_deviates(w.seed, nl * nyi * nxi).reshape(nl, nyi, nxi)
    q[:, 1:-1, 1:-1] += dq
    return q


def _dst1(x: np.ndarray, axis: int) -> np.ndarray:
    """FFTW RODFT00 (unnormalised) along `axis`, via the odd extension."""
    x = np.moveaxis(x, axis, -1)
    n = x.shape[-1]
    z = np.zeros(x.shape[:-1] + (2 * (n + 1),))
    z[..., 1:n + 1] = x
    z[..., n + 2:] = -x[..., ::-1]
    Z = np.fft.fft(z, axis=-1)
    return np.moveaxis(-Z[..., 1:n + 1].imag, -1, axis)


def initial_p(w: Workload) -> np.ndarray:
    rng = np.random.default_rng(w.seed)
    nl, nyp, nxp = w.nl, w.nyp, w.nxp
    if w.cyclic:
        nyi, nxi = nyp - 2, nxp - 1
    else:
        nyi, nxi = nyp - 2, nxp - 2
    p = np.zeros((nl, nyp, nxp))
    for k in range(nl):
        dq = rng.standard_normal((nyi, nxi))
        if w.init == "k3":
            # shape the field in sine space: amplitude ~ kappa^-2 (1-D spectrum ~ k^-3)
            ky = np.pi * np.arange(1, nyi + 1) / ((nyi + 1) * w.dx)
            kx = np.pi * np.arange(1, nxi + 1) / ((nxi + 1) * w.dx)
            kap2 = ky[:, None] ** 2 + kx[None, :] ** 2
            dq = _dst1(_dst1(dq, 0), 1) / kap2
            dq = _dst1(_dst1(dq, 0), 1)
        dq *= w.q_rms / np.sqrt(np.mean(dq * dq))
        # invert Del^2 p / f0 = dq with p = 0 on the boundary (periodic in x for a channel)
        if w.cyclic:
            D = np.fft.rfft(_dst1(dq, 0), axis=1)
            ly = 2.0 * (np.cos(np.pi * np.arange(1, nyi + 1) / (nyi + 1)) - 1.0) / w.dx ** 2
            lx = 2.0 * (np.cos(2.0 * np.pi * np.arange(nxi // 2 + 1) / nxi) - 1.0) / w.dx ** 2
            P = w.fnot * D / (ly[:, None] + lx[None, :])
            pin = _dst1(np.fft.irfft(P, n=nxi, axis=1), 0) / (2.0 * (nyi + 1))
            p[k, 1:-1, :-1] = pin
            p[k, 1:-1, -1] = pin[:, 0]
        else:
            D = _dst1(_dst1(dq, 0), 1)
            ly = 2.0 * (np.cos(np.pi * np.arange(1, nyi + 1) / (nyi + 1)) - 1.0) / w.dx ** 2
            lx = 2.0 * (np.cos(np.pi * np.arange(1, nxi + 1) / (nxi + 1)) - 1.0) / w.dx ** 2
            P = w.fnot * D / (ly[:, None] + lx[None, :])
            p[k, 1:-1, 1:-1] = _dst1(_dst1(P, 0), 1) / (4.0 * (nyi + 1) * (nxi + 1))
    # rescale so the interior PV anomaly Del^2 p / f0 - f0 A p has rms q_rms
    A = np.zeros((nl, nl))
    for k in range(nl):
        if k > 0:
            A[k, k - 1] = -1.0 / (w.gp[k - 1] * w.h[k])
        if k < nl - 1:
            A[k, k + 1] = -1.0 / (w.gp[k] * w.h[k])
        A[k, k] = -A[k].sum()
    lap = (p[:, 1:-1, :-2] + p[:, 1:-1, 2:] + p[:, :-2, 1:-1] + p[:, 2:, 1:-1] - 4.0 * p[:, 1:-1, 1:-1]) / w.dx ** 2
    qa = lap / w.fnot - w.fnot * np.einsum("kl,lij->kij", A, p[:, 1:-1, 1:-1])
    p *= w.q_rms / np.sqrt(np.mean(qa * qa))
    return p


def ocean_box(name: str, n: int, dx: float, ah4: float, dt: float, days: float, nl: int = 3,
              **kw) -> Workload:
    lay = OCEAN_LAYERS[nl]
    return Workload(name=name, nxp=n, nyp=n, nl=nl, dx=dx, dt=dt, days=days, h=lay["h"], gp=lay["gp"],
                    ah2=[0.0] * nl, ah4=[ah4] * nl, **kw)


def c1_box129() -> Workload:
    """129^2 box at 30 km, ah4 = 2e10 m^4/s, dt = 1 h, 10 days; q = beta yprel +
    1e-7 N(0, 1) PV noise from SplitMix64 (seed 1), p by one inversion."""
    return ocean_box("box129_dx30km_3l", 129, 30.0e3, 2.0e10, 3600.0, 10.0, init="splitmix", q_rms=1.0e-7, seed=1)


def c2_box513() -> Workload:
    """513^2 box at 7.5 km, ah4 = 1e10 m^4/s, dt = 20 min, 30 days; k^-3 PV field rms 5e-6 (seed 2)."""
    return ocean_box("box513_dx7.5km_3l", 513, 7.5e3, 1.0e10, 1200.0, 30.0, init="k3", q_rms=5.0e-6, seed=2)


def c5_box257() -> Workload:
    """257^2 box at 15 km, ah4 = 1e10 m^4/s, dt = 20 min, 30 days; k^-3 PV field rms 5e-6 (seed 5)."""
    return ocean_box("box257_dx15km_3l", 257, 15.0e3, 1.0e10, 1200.0, 30.0, init="k3", q_rms=5.0e-6, seed=5)


def ocean385() -> Workload:
    """385^2 ocean box at 10 km (the coupled configuration's ocean), dt = 20 min."""
    return ocean_box("box385_dx10km_3l", 385, 10.0e3, 1.0e10, 1200.0, 30.0, init="k3", q_rms=5.0e-6, seed=4)


def atmos_channel() -> Workload:
    """97 x 49 zonally periodic channel at 120 km (11520 x 5760 km), H = 2000/3000/4000 m,
    g' = 1.2/0.4, ah4 = 2e13, bcco = 1, delek = 0, dz_sign = +1, tau_sign = -1; dt = 200 s."""
    return Workload(name="atmos97x49_dx120km_3l", nxp=97, nyp=49, nl=3, dx=120.0e3, dt=200.0, days=30.0,
                    cyclic=True, h=[2000.0, 3000.0, 4000.0], gp=[1.2, 0.4], ah2=[0.0] * 3, ah4=[2.0e13] * 3,
                    bcco=1.0, delek=0.0, dz_sign=1, tau_sign=-1, tau0=0.1, init="k3", q_rms=5.0e-6, seed=6,
                    supp_every=100)


ENS_SIZES = (65, 129, 257, 385, 513)
BASIN_L = 3840.0e3   # square basin side (m)


def stable_dt(dx: float, ah4: float, cap: float = 1200.0) -> float:
    """Largest divisor of one day (s) not above min(cap, 0.5 / (64 ah4 / dx^4)),
    the biharmonic stability bound of the leapfrog step."""
    lim = min(cap, 0.5 / (64.0 * ah4 / dx ** 4))
    return float(max(d for d in range(1, 86401) if 86400 % d == 0 and d <= lim))


def ensemble_member(i: int, base_seed: int = 3, days: float = 10.0) -> Workload:
    """Member i of the uneven double-gyre ensemble (seed base_seed + i): grid
    drawn from 65^2..513^2 over the same 3840 km basin, 2-4 layers, wind
    amplitude tau0 ~ U(0.05, 0.15) N/m^2, biharmonic viscosity ~ logU(1e9, 1e11)
    m^4/s, PV noise 1e-7 N(0, 1) 1/s from SplitMix64 (seed base_seed + i) with
    p by one inversion (as the 129^2 box), the stable time step for its grid."""
    rng = np.random.default_rng(base_seed + i)
    n = int(rng.choice(ENS_SIZES))
    nl = int(rng.choice([2, 3, 4]))
    tau0 = float(rng.uniform(0.05, 0.15))
    ah4 = float(10.0 ** rng.uniform(9.0, 11.0))
    dx = BASIN_L / (n - 1)
    return ocean_box(f"ens{i:02d}_box{n}_{nl}l", n, dx, ah4, stable_dt(dx, ah4), days, nl=nl, init="splitmix",
                     q_rms=1.0e-7, seed=base_seed + i, tau0=tau0)


def uneven_ensemble(n_members: int = 64, base_seed: int = 3, days: float = 10.0) -> list:
    """The 64-member batched-ensemble workload."""
    return [ensemble_member(i, base_seed, days) for i in range(n_members)]


def small_box(n: int = 33, nl: int = 3, seed: int = 7) -> Workload:
    """A small box for fast parity tests (30 km spacing)."""
    return ocean_box(f"box{n}_{nl}l", n, 30.0e3, 2.0e10, 3600.0, 1.0, nl=nl, init="k3", q_rms=5.0e-6, seed=seed)


def small_channel(nx: int = 33, ny: int = 17, nl: int = 3, seed: int = 8) -> Workload:
    w = atmos_channel()
    w.name, w.nxp, w.nyp, w.nl, w.seed = f"chan{nx}x{ny}_{nl}l", nx, ny, nl, seed
    if nl != 3:
        lay = {2: ([3000.0, 6000.0], [0.8]), 4: ([1500.0, 2000.0, 2500.0, 3000.0], [1.2, 0.8, 0.4])}[nl]
        w.h, w.gp, w.ah2, w.ah4 = lay[0], lay[1], [0.0] * nl, [2.0e13] * nl
    return w


WORKLOADS = {
    "box129": c1_box129,
    "box513": c2_box513,
    "box257": c5_box257,
    "box385": ocean385,
    "atmos97x49": atmos_channel,
}
