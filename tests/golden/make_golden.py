"""Generate the golden fixtures in tests/golden/golden.npz with numpy/scipy.

These pin the oracle's building blocks against library routines, not against
the oracle itself:
  dst1_*      scipy.fft.dst(type=1): FFTW's RODFT00 definition, unnormalised
  rfft_*      numpy.fft.rfft: FFTW's r2c definition
  thomas_*    scipy.linalg.solve_banded on the y-tridiagonal of inhomog.F90:188-199
  helm_*      dense numpy.linalg.solve of the 5-point Dirichlet (box) or
              x-periodic (channel) operator minus rdm2 on tiny grids
  eig_*       numpy.linalg.eig of the layer-coupling matrix A (modes.F90:221-246)
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.fft
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))


def A_matrix(h, gp):
    nl = len(h)
    A = np.zeros((nl, nl))
    for k in range(nl):
        if k > 0:
            A[k, k - 1] = -1.0 / (gp[k - 1] * h[k])
        if k < nl - 1:
            A[k, k + 1] = -1.0 / (gp[k] * h[k])
        A[k, k] = -A[k].sum()
    return A


def helm_dense(nxp, nyp, dx, dy, rdm2, cyclic, rhs):
    """Solve (d2x + d2y - rdm2) u = rhs on the interior with u = 0 on solid walls."""
    if cyclic:
        xs = list(range(nxp - 1))
    else:
        xs = list(range(1, nxp - 1))
    ys = list(range(1, nyp - 1))
    idx = {(j, i): n for n, (j, i) in enumerate((j, i) for j in ys for i in xs)}
    N = len(idx)
    M = np.zeros((N, N))
    b = np.zeros(N)
    for (j, i), n in idx.items():
        M[n, n] = -2.0 / dx ** 2 - 2.0 / dy ** 2 - rdm2
        for di in (-1, 1):
            ii = i + di
            if cyclic:
                ii %= (nxp - 1)
            if (j, ii) in idx:
                M[n, idx[(j, ii)]] += 1.0 / dx ** 2
        for dj in (-1, 1):
            if (j + dj, i) in idx:
                M[n, idx[(j + dj, i)]] += 1.0 / dy ** 2
        b[n] = rhs[j, i]
    u = np.linalg.solve(M, b)
    out = np.zeros((nyp, nxp))
    for (j, i), n in idx.items():
        out[j, i] = u[n]
    if cyclic:
        out[:, -1] = out[:, 0]
    return out


def main():
    rng = np.random.default_rng(20261203)
    g = {}
    for n in (63, 127, 255, 383, 511):
        x = rng.standard_normal(n)
        g[f"dst1_in_{n}"] = x
        g[f"dst1_out_{n}"] = scipy.fft.dst(x, type=1)
    for n in (96, 64):
        x = rng.standard_normal(n)
        g[f"rfft_in_{n}"] = x
        g[f"rfft_out_{n}"] = np.fft.rfft(x)
    # y-tridiagonal systems a u[j-1] + b u[j] + a u[j+1] = f[j] (box wavenumbers of a 129^2 grid)
    dx = dy = 30.0e3
    a = 1.0 / dy ** 2
    nyi = 127
    for k, rdm2 in ((1, 0.0), (7, 0.0), (64, 1.0 / 30e3 ** 2), (127, 1.0 / 50e3 ** 2)):
        b = -2.0 * a + 2.0 / dx ** 2 * (np.cos(k * 2 * np.pi / 256) - 1.0) - rdm2
        f = rng.standard_normal(nyi)
        ab = np.zeros((3, nyi))
        ab[0, 1:] = a
        ab[1, :] = b
        ab[2, :-1] = a
        g[f"thomas_k{k}_ab"] = np.array([a, b])
        g[f"thomas_k{k}_f"] = f
        g[f"thomas_k{k}_u"] = scipy.linalg.solve_banded((1, 1), ab, f)
    # tiny Helmholtz problems, rdm2 = f0^2 |eig(A)| of the 3-layer ocean (f0 = 1e-4), numpy eig
    w3 = np.sort(np.abs(np.linalg.eig(A_matrix([300.0, 1100.0, 2600.0], [0.05, 0.025]))[0].real))
    rdm2_3 = (0.0, 1e-8 * w3[1], 1e-8 * w3[2])
    for name, (nxp, nyp, cyc) in {"box9x7": (9, 7, False), "chan17x9": (17, 9, True), "box17x9": (17, 9, False)}.items():
        rhs = rng.standard_normal((nyp, nxp))
        if cyc:
            rhs[:, -1] = rhs[:, 0]
        for m, rd in enumerate(rdm2_3):
            g[f"helm_{name}_rhs"] = rhs
            g[f"helm_{name}_rdm2_{m}"] = np.array([rd])
            g[f"helm_{name}_u_{m}"] = helm_dense(nxp, nyp, 10e3, 10e3, rd, cyc, rhs)
    # layer-coupling eigenproblems
    sets = {
        "ocean3": ([300.0, 1100.0, 2600.0], [0.05, 0.025]),
        "ocean2": ([1000.0, 3000.0], [0.02]),
        "ocean4": ([300.0, 700.0, 1200.0, 1800.0], [0.05, 0.03, 0.02]),
        "atmos3": ([2000.0, 3000.0, 4000.0], [1.2, 0.4]),
    }
    for name, (h, gp) in sets.items():
        A = A_matrix(h, gp)
        w, v = np.linalg.eig(A)
        order = np.argsort(np.abs(w), kind="stable")
        g[f"eig_{name}_h"] = np.array(h)
        g[f"eig_{name}_gp"] = np.array(gp)
        g[f"eig_{name}_A"] = A
        g[f"eig_{name}_w"] = np.abs(w.real[order])
        g[f"eig_{name}_v"] = v.real[:, order]
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **g)
    print(f"wrote {len(g)} arrays")


if __name__ == "__main__":
    main()
