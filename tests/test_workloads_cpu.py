"""The synthetic initial states: SplitMix64 against its published test
vector, the Box-Muller deviates' moments and the layout of the 129^2 box's
initial PV; the oracle's start by one inversion against the closed form of
a single Helmholtz mode."""
import numpy as np

from openqg_amd import workloads as W


def test_splitmix64_test_vector():
    # the first outputs of the reference SplitMix64 from state 1234567
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [int(v) for v in W.splitmix64(1234567, 5)] == want


def test_normal_deviates_moments_and_pairs():
    x = W.normal_deviates(1, 400000)
    assert abs(x.mean()) < 4.0 / np.sqrt(x.size)
    assert abs(x.std() - 1.0) < 4.0 / np.sqrt(2 * x.size)
    # the fourth moment of a standard normal is 3
    assert abs(np.mean(x ** 4) - 3.0) < 0.05
    # pairs share their radius: z0^2 + z1^2 = -2 ln(1 - u1)
    z = W.splitmix64(1, 2)
    u1 = float(z[0] >> np.uint64(11)) * 2.0 ** -53
    assert np.isclose(x[0] ** 2 + x[1] ** 2, -2.0 * np.log1p(-u1), rtol=1e-14)
    # an odd count is the prefix of the even one
    assert np.array_equal(W.normal_deviates(7, 5), W.normal_deviates(7, 6)[:5])


def test_box129_initial_pv_layout():
    w = W.c1_box129()
    assert w.init == "splitmix" and w.seed == 1
    q = W.initial_q(w)
    nl, nyp, nxp = q.shape
    yprel = -0.5 * (nyp - 1) * w.dx + np.arange(nyp) * w.dx
    dq = q - (w.beta * yprel)[None, :, None]
    assert np.all(dq[:, 0, :] == 0) and np.all(dq[:, -1, :] == 0)
    assert np.all(dq[:, :, 0] == 0) and np.all(dq[:, :, -1] == 0)
    d = W.normal_deviates(1, 3)
    # interior points in order x, then y, then layers
    assert np.allclose(dq[0, 1, 1:4], 1e-7 * d, rtol=0, atol=1e-20)
    n = W.normal_deviates(1, nl * (nyp - 2) * (nxp - 2))
    assert np.allclose(dq[2, 1:-1, 1:-1].ravel(), 1e-7 * n[2 * (nyp - 2) * (nxp - 2):], rtol=0, atol=1e-20)


def test_oracle_start_by_one_inversion():
    """One Helmholtz eigenmode: a barotropic q - beta yprel = A sin(pi i/(nx-1))
    sin(pi j/(ny-1)) in every layer inverts to p = f0 A s / lambda with the
    5-point eigenvalue lambda (the barotropic mode has no stretching and the
    mass correction with dpi = 0 leaves it alone: the layers are equal)."""
    from oracle import cpuref as cr

    w = W.small_box(33, 3)
    o = cr.Basin(w.params())
    nyp, nxp = w.nyp, w.nxp
    i = np.arange(nxp)[None, :]
    j = np.arange(nyp)[:, None]
    s = np.sin(np.pi * i / (nxp - 1)) * np.sin(np.pi * j / (nyp - 1))
    A = 1e-6
    yprel = -0.5 * (nyp - 1) * w.dx + np.arange(nyp) * w.dx
    q0 = np.broadcast_to((w.beta * yprel)[:, None] + A * s, (w.nl, nyp, nxp)).copy()
    o.start_from_q(q0)
    q, qm, p, pm = o.get_state()
    lam = 2.0 * (np.cos(np.pi / (nxp - 1)) - 1.0) / w.dx ** 2 + 2.0 * (np.cos(np.pi / (nyp - 1)) - 1.0) / w.dx ** 2
    want = w.fnot * A * s / lam
    for k in range(w.nl):
        assert np.abs(p[k] - want).max() <= 1e-10 * np.abs(want).max(), k
    assert np.array_equal(q, qm) and np.array_equal(p, pm)
    assert np.allclose(q[:, 1:-1, 1:-1], q0[:, 1:-1, 1:-1], rtol=0, atol=1e-20)
