#!/usr/bin/env python3
"""Benchmark: simulated days per second of the 513^2 3-layer FP64 ocean box.

One "step" is one QG timestep (step_qg) of the 513 x 513 x 3 double-gyre box
at dx = 7.5 km, dt = 20 min, with the leapfrog suppression every 25 steps, on
synthetic seeded input resident in HBM.  The default K = 2160 steps is the
full 30-day run.  Prints one JSON line.

Also reported:
  roofline      the dominant kernel's algorithmic bytes per launch / its mean
                duration (HIP events on the library's stream), against 8 TB/s;
  cpu_baseline  the CPU oracle (a C restatement of the reference step) on the
                host cores, on a bounded sample of the same workload;
  ensemble      64 uneven basins (65^2..513^2, 2-4 layers) for 10 simulated
                days each: batched launches against the same members run one
                after another by the plain path (aggregate simulated days/s);
  adjoint       gradient of a kinetic-energy objective after 30 days of the
                257^2 3-layer box (forward + hand-written adjoint) against the
                plain forward run, store-all and with 32 revolve snapshots;
                and a batch of 16 perturbed members' gradients, one after
                another and concurrently (one context and stream each); the
                same gradient of the 97 x 49 atmosphere channel;
  coupled       the coupled configuration for 30 days: the 385^2 ocean with its
                100 m mixed layer and the 97 x 49 atmosphere with its 1000 m
                mixed layer, exchanging stress, heat flux and entrainment every
                ocean step (dynamic_step in one captured graph per cadence
                period), against the two basins' QG-only runs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNEL_REPS = 20   # back-to-back launches per kernel in the event timing
METRIC = "simulated days/s, 513² 3-layer ocean FP64; batched ensemble + adjoint vs plain"


def algorithmic_bytes(w, kernel: str) -> float:
    """Compulsory HBM bytes of one launch of `kernel` (F = one layer field)."""
    F = 8.0 * w.nxp * w.nyp
    nl = w.nl
    return {
        "k_vort": (5 * nl + 2) * F,      # read q, p, pm, qm, ent1, wekp; write q_new
        "k_fwd": 2 * nl * F,             # read q_new; write x spectra
        "k_ysolve": 2 * nl * F,          # read + write spectra
        "k_inv": (3 * nl - 1) * F,       # read spectra + hom solutions; write p_new
        "step": (6 * nl + 2) * F,        # whole step: q, qm, p, pm, wekp, ent1 in; q, p out
    }[kernel]


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_seconds(w, threads: int, nsteps: int, reps: int = 3) -> float:
    """Median over `reps` samples of the oracle's time for `nsteps` whole steps of
    workload w (each sample from the same initial state, after one untimed step)."""
    from oracle import cpuref as cr

    from openqg_amd import workloads as W

    cr.set_threads(threads)
    ent, wekp, txis, txin = w.forcing()
    p = None if w.init == "splitmix" else w.initial_p()
    times = []
    for _ in range(reps):
        b = cr.Basin(w.params())
        if p is None:
            b.start_from_q(W.initial_q(w))
        else:
            b.set_state(p=p, pm=p)
            b.init_state()
        b.step(w.tdt, ent, wekp, txis, txin)   # first step (page-in, thread start) untimed
        b.suppress()
        t0 = time.perf_counter()
        for n in range(nsteps):
            nt = n + 2
            b.step(w.tdt, ent, wekp, txis, txin)
            if (nt - 1) % w.supp_every == 0:
                b.suppress()
        times.append(time.perf_counter() - t0)
    return float(sorted(times)[len(times) // 2])


def cpu_baseline(w, seconds: float):
    """The CPU oracle (a restatement of the reference step, not the reference
    binary) timed on this host: median of 3 samples of a fixed step count sized
    to about `seconds`/10 each, with the OpenMP threads of this process's CPU
    share (OMP_NUM_THREADS, which the GPU hosts set to the cores allotted per
    GPU; otherwise every visible core) and with 1 thread; plus the 129^2 box
    and the coupled 385^2 + 97x49 pair with the same threads."""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    cores = int(env) if env.isdigit() and int(env) > 0 else visible
    cores = max(1, min(cores, visible))

    def days_per_s(wk, threads, nsteps):
        el = _oracle_seconds(wk, threads, nsteps)
        return nsteps * wk.dt / 86400.0 / el, el

    # size the samples from one short probe
    probe = _oracle_seconds(w, cores, 10, reps=1) / 10
    n = max(20, int(seconds / 10 / probe))
    v, el = days_per_s(w, cores, n)
    probe1 = _oracle_seconds(w, 1, 3, reps=1) / 3
    n1 = max(5, int(seconds / 20 / probe1))
    v1, el1 = days_per_s(w, 1, n1)
    from openqg_amd import workloads as W

    w1 = W.c1_box129()
    v129, el129 = days_per_s(w1, cores, w1.nsteps)
    # coupled pair, one simulated day: 432 atmosphere steps of 200 s, 72 ocean steps of 1200 s
    wo, wa = W.ocean385(), W.atmos_channel()
    _, elo = days_per_s(wo, cores, 72)
    _, ela = days_per_s(wa, cores, 432)
    return {
        "value": v,
        "unit": "simulated days/s",
        "cores": cores,
        "kind": "port",
        "sample": f"median of 3 runs of {n} steps of {w.name} (dt={w.dt:g} s) after init, OpenMP {cores} threads, "
                  f"{el:.2f} s each",
        "cpu_model": _cpu_model(),
        "visible_cores": visible,
        "threads_reason": "OMP_NUM_THREADS (the host's CPU share per GPU)" if env.isdigit() else "all visible cores",
        "value_1thread": v1,
        "sample_1thread": f"median of 3 runs of {n1} steps, 1 thread, {el1:.2f} s each",
        "box129": {"value": v129, "unit": "simulated days/s", "sample": f"median of 3 runs of the whole {w1.nsteps}-step "
                   f"10-day run (dt={w1.dt:g} s), {cores} threads, {el129:.2f} s each"},
        "coupled_385_97x49": {"value": 1.0 / (elo + ela), "unit": "simulated days/s",
                              "sample": f"one day: 72 ocean steps of 385^2 ({elo:.2f} s) + 432 atmosphere steps of "
                                        f"97x49 ({ela:.2f} s), medians of 3, {cores} threads, QG only"},
    }


def ensemble_bench(members: int, days: float):
    """The uneven ensemble (grids 65^2..513^2, 2-4 layers, seeded wind, viscosity
    and PV noise; each member its stable time step) for `days` simulated days
    per member: batched (oqg_ens_*, launch schedule picked by the tuner; all
    members in one batch, and the 513^2 members on their own streams beside
    the rest in two batches of equal work on two streams, the faster of the
    two reported) against
    the same members advanced one after another by the plain graph path, each
    finished before the next starts; also against the plain runs all enqueued
    at once (each member on its own stream, overlapping on the GPU) and
    against one eager step_qg call per member step (a 25-step sample of each
    member, scaled to its step count).  All from the same initial states,
    resident forcing, suppression every 25 steps."""
    from openqg_amd import workloads as W
    from openqg_amd.ensemble import Ensemble, EnsembleGroups

    ws = W.uneven_ensemble(members, days=days)

    def make():
        qgs = []
        for w in ws:
            qg = w.make_qg()
            w.start(qg)
            ent, wekp, txis, txin = w.forcing()
            qg.set_forcing(ent, wekp, txis, txin)
            qg.sync()
            qgs.append(qg)
        return qgs

    qgs = make()
    tdt = [w.tdt for w in ws]
    nsteps = [w.nsteps for w in ws]
    sim_days = sum(n * w.dt for n, w in zip(nsteps, ws)) / 86400.0
    # plain path, one member after another, each finished before the next
    # starts (each warmed up: graph captured)
    for w, qg in zip(ws, qgs):
        qg.run(w.tdt, 2 * w.supp_every, 0, w.supp_every)
        qg.sync()
    t0 = time.perf_counter()
    for w, qg, n in zip(ws, qgs, nsteps):
        qg.run(w.tdt, n, 2 * w.supp_every, w.supp_every)
        qg.sync()
    seq = time.perf_counter() - t0
    # the same runs enqueued without waiting: every member on its own stream,
    # so the GPU overlaps them as it can
    t0 = time.perf_counter()
    for w, qg, n in zip(ws, qgs, nsteps):
        qg.run(w.tdt, n, 2 * w.supp_every + n, w.supp_every)
    for qg in qgs:
        qg.sync()
    conc = time.perf_counter() - t0
    # the drop-in call pattern: one step_qg call per member step (eager
    # launches), timed on a 25-step sample of every member and scaled
    sample = 25
    eager = 0.0
    for w, qg, n in zip(ws, qgs, nsteps):
        t0 = time.perf_counter()
        for nt in range(1, sample + 1):
            qg.step(w.tdt)
            if (nt - 1) % w.supp_every == 0:
                qg.suppress()
        qg.sync()
        eager += (time.perf_counter() - t0) * n / sample
    # batched: tune the launch schedule, warm up (plans and graphs), time
    ens = Ensemble(qgs)
    sched, tune_ms = ens.tune(tdt, trial_steps=2 * 25, supp_every=25)
    # warm-up: the uneven run itself, twice, so that the launch plans of every
    # segment (members drop out as they finish) exist for both leapfrog parities
    for _ in range(2):
        ens.run(tdt, nsteps, 0, 25)
    ens.sync()
    t0 = time.perf_counter()
    ens.run(tdt, nsteps, 2 * 25, 25)
    ens.sync()
    bat1 = time.perf_counter() - t0
    ens.close()
    # the 513^2 members on their own streams (plain graph replay), the rest
    # batched in two groups of equal work per step, each on its own stream
    grp = EnsembleGroups(qgs, 2, solo_points=513 * 513)
    grp.tune(tdt, trial_steps=2 * 25, supp_every=25)
    for _ in range(2):
        grp.run(tdt, nsteps, 0, 25)
    grp.sync()
    t0 = time.perf_counter()
    grp.run(tdt, nsteps, 2 * 25, 25)
    grp.sync()
    bat2 = time.perf_counter() - t0
    grp.close()
    bat = min(bat1, bat2)
    for qg in qgs:
        qg.close()
    return {
        "workload": f"uneven_ensemble_{members}x{days:g}d",
        "members": members,
        "grids": sorted({f"{w.nxp}^2x{w.nl}" for w in ws}),
        "member_steps": int(sum(nsteps)),
        "value": round(sim_days / bat, 2),
        "sequential_value": round(sim_days / seq, 2),
        "concurrent_streams_value": round(sim_days / conc, 2),
        "sequential_per_call_value": round(sim_days / eager, 2),
        "unit": "aggregate simulated days/s",
        "speedup_vs_sequential": round(seq / bat, 2),
        "speedup_vs_concurrent_streams": round(conc / bat, 2),
        "speedup_vs_sequential_per_call": round(eager / bat, 2),
        "batched_s": round(bat, 4),
        "batched_one_group_s": round(bat1, 4),
        "batched_solo_513_two_groups_s": round(bat2, 4),
        "mode": "solo_513_plus_two_groups" if bat2 < bat1 else "one_group",
        "sequential_s": round(seq, 4),
        "concurrent_streams_s": round(conc, 4),
        "sequential_per_call_s": round(eager, 4),
        "schedule": sched,
        "tune_ms": [round(t, 3) for t in tune_ms],
    }


def progress(msg: str):
    """One line per finished leg on stderr (stdout carries the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def adjoint_bench(days: float, snapshots: int, batch: int = 16):
    """Gradient of the layer-1 kinetic energy after `days` of the 257^2 3-layer
    box with respect to the wind stress and the initial PV: forward + adjoint
    (store-all snapshots, and binomial checkpointing with `snapshots` slots)
    against the plain forward run of the same steps (graph replay); the same
    for the 97 x 49 atmosphere channel (dt 200 s) under the `channel` key."""
    from openqg_amd import workloads as W

    w = W.c5_box257()
    n = int(round(days * 86400.0 / w.dt))
    out = _gradient_timing(w, n, snapshots)
    out["workload"] = f"{w.name}_{days:g}d_ke_gradient"
    progress(f"adjoint {w.name}: revolve {out[f'revolve_{snapshots}']['ratio_vs_forward']}x the forward run")
    if batch > 1:
        out["batch"] = adjoint_batch_bench(w, n, snapshots, batch)
        progress(f"adjoint batch: {out['batch']['s']} s")
    # the atmosphere channel (97 x 49, dt 200 s) over the same days
    wc = W.atmos_channel()
    nc = int(round(days * 86400.0 / wc.dt))
    out["channel"] = dict(workload=f"{wc.name}_{days:g}d_ke_gradient", **_gradient_timing(wc, nc, snapshots))
    progress(f"adjoint {wc.name}: store-all {out['channel']['store_all']['ratio_vs_forward']}x the forward run")
    return out


def _gradient_timing(w, n: int, snapshots: int):
    """Forward run and gradients (store-all, revolve) of workload w over n steps."""
    from openqg_amd.adjoint import gradient
    from openqg_amd.forcing import double_gyre_tau

    qg = w.make_qg()
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    q0 = qg.get_state()[0]
    taux, tauy = double_gyre_tau(w.nyp, w.nxp, w.dx, w.tau0)
    ent, wekp, txis, txin = w.forcing()
    qg.set_forcing(ent, wekp, txis, txin)
    # plain forward (graphs captured in the warm-up)
    qg.run(w.tdt, 2 * w.supp_every, 0, w.supp_every)
    qg.sync()
    t0 = time.perf_counter()
    qg.run(w.tdt, n, 0, w.supp_every)
    qg.sync()
    fwd = time.perf_counter() - t0
    out = {"steps": n, "forward_s": round(fwd, 4)}
    for name, snaps in (("store_all", None), (f"revolve_{snapshots}", snapshots)):
        gradient(qg, q0, taux, tauy, w.tdt, n, w.supp_every, snaps)   # allocation, graph capture
        qg.sync()
        t0 = time.perf_counter()
        J, gx, gy, gq, fsteps = gradient(qg, q0, taux, tauy, w.tdt, n, w.supp_every, snaps)
        el = time.perf_counter() - t0
        out[name] = {"s": round(el, 4), "ratio_vs_forward": round(el / fwd, 3), "forward_steps": int(fsteps)}
    out["J"] = J
    qg.close()
    return out


def coupled_bench(days: float):
    """The coupled configuration: the 385^2 ocean box (dt = 1200 s) with a
    100 m mixed layer and the 97 x 49 atmosphere channel (dt = 200 s, nstr = 6)
    with a 1000 m mixed layer, coupled through the device coupler
    (openqg_amd.coupled: stress, heat flux, both mixed layers, suppression every
    25 ocean and 100 atmosphere steps), radiation coefficients from the
    double-gyre radiative balance, mixed layers started at radiative balance;
    `days` simulated days after a one-day warm-up (graph capture).  Against the
    two basins' QG-only runs one after the other."""
    from dataclasses import replace

    import numpy as np

    from openqg_amd import workloads as W
    from openqg_amd.coupled import Coupler, load_grid
    from openqg_amd.mixed import OceanMixedLayer, zonal_temp
    from openqg_amd.radiation import coupled_setup

    wo, wa = W.ocean385(), W.atmos_channel()
    nstr = int(round(wo.dt / wa.dt))

    def make(w):
        qg = w.make_qg()
        p = w.initial_p()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        qg.set_forcing(*w.forcing())
        return qg

    o, a = make(wo), make(wa)
    go, ga = replace(o.b, hm=100.0), replace(a.b, hm=1000.0)
    g = load_grid(go, ga)
    prm, octemp, attemp, rad, xc1, _ = coupled_setup(go, ga, yo0=(g.ny1 - 1) * ga.dx)
    oml = OceanMixedLayer(o, octemp, fluxn=False, fluxs=True, st2d=380.0, st4d=4.0e10, ycexp=1.0, hm=go.hm)
    sst = np.array([zonal_temp(go, octemp, j) for j in range(1, go.nyt + 1)])[:, None] + np.zeros((1, go.nxt))
    oml.set_state(sst, sst)
    cp = Coupler(oml, a, prm, xc1ast=xc1)
    ast = np.array([zonal_temp(ga, attemp, j) for j in range(1, ga.nyt + 1)])[:, None] + np.zeros((1, ga.nxt))
    h = np.full(ast.shape, ga.hm)
    cp.set_aml_state(ast, ast, h, h)
    cp.xforc()     # the initial exchange (openqg.F90:456-478)
    cp.forcing()
    nday = int(round(86400.0 / wa.dt))
    cp.run(nday, 0, wo.tdt, wa.tdt, nstr)   # one day: captures the cadence graph
    o.sync()
    na = int(round(days * nday))
    t0 = time.perf_counter()
    cp.run(na, nday, wo.tdt, wa.tdt, nstr)
    o.sync()
    t_cpl = time.perf_counter() - t0
    bad = o.diagnose(wo.dt)["nonfinite"] + a.diagnose(wa.dt)["nonfinite"]
    s_o, s_m = oml.get_state()
    ast1, _, hm1, _ = cp.get_aml_state()
    bad += float(np.count_nonzero(~np.isfinite(s_o)) + np.count_nonzero(~np.isfinite(ast1)))
    cp.close()
    oml.close()
    no = na // nstr
    # the same basins QG-only, one after the other
    t0 = time.perf_counter()
    o.run(wo.tdt, no, 0, 25)
    o.sync()
    t_o = time.perf_counter() - t0
    t0 = time.perf_counter()
    a.run(wa.tdt, na, 0, 100)
    a.sync()
    t_a = time.perf_counter() - t0
    o.close()
    a.close()
    return {"workload": f"coupled_ocean385_oml100m_atmos97x49_aml1000m_{days:g}d", "ocean_steps": no,
            "atmos_steps": na, "value": round(days / t_cpl, 2), "unit": "simulated days/s", "s": round(t_cpl, 4),
            "nonfinite_after": bad, "sst_range": [round(float(s_o.min()), 3), round(float(s_o.max()), 3)],
            "ast_range": [round(float(ast1.min()), 3), round(float(ast1.max()), 3)],
            "hmixa_range": [round(float(hm1.min()), 1), round(float(hm1.max()), 1)],
            "qg_only_ocean_s": round(t_o, 4), "qg_only_atmos_s": round(t_a, 4),
            "qg_only_one_after_other_value": round(days / (t_o + t_a), 2)}


def adjoint_batch_bench(w, n: int, snapshots: int, members: int):
    """The gradient of `members` perturbed members (initial PV seed and wind
    amplitude vary), each on its own context and stream with binomial
    checkpointing: one after another, then concurrently from host threads."""
    import dataclasses

    from openqg_amd.adjoint import gradient_batch
    from openqg_amd.forcing import double_gyre_tau

    qgs, q0s, txs, tys = [], [], [], []
    for i in range(members):
        wi = dataclasses.replace(w, seed=w.seed + i, tau0=w.tau0 * (0.9 + 0.2 * i / max(members - 1, 1)))
        qg = wi.make_qg()
        p = wi.initial_p()
        qg.set_state(p=p, pm=p)
        qg.init_state()
        q0s.append(qg.get_state()[0])
        tx, ty = double_gyre_tau(wi.nyp, wi.nxp, wi.dx, wi.tau0)
        txs.append(tx)
        tys.append(ty)
        qgs.append(qg)
    gradient_batch(qgs, q0s, txs, tys, w.tdt, n, w.supp_every, snapshots)   # allocation, graph capture
    t0 = time.perf_counter()
    seq = gradient_batch(qgs, q0s, txs, tys, w.tdt, n, w.supp_every, snapshots, concurrent=False)
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    con = gradient_batch(qgs, q0s, txs, tys, w.tdt, n, w.supp_every, snapshots)
    t_con = time.perf_counter() - t0
    import numpy as np

    same = all(a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3]) for a, b in zip(seq, con))
    for qg in qgs:
        qg.close()
    return {"members": members, "snapshots": snapshots, "s": round(t_con, 4), "sequential_s": round(t_seq, 4),
            "speedup_vs_sequential": round(t_seq / t_con, 3), "identical_to_sequential": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2160)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="box513")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--flags", type=int, default=0, help="oqg_desc.flags (1: per-thread Thomas y-solve)")
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of replaying graphs")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ens-members", type=int, default=64)
    ap.add_argument("--ens-days", type=float, default=10.0)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint measurement")
    ap.add_argument("--adj-days", type=float, default=30.0)
    ap.add_argument("--adj-snapshots", type=int, default=32)
    ap.add_argument("--adj-batch", type=int, default=16, help="members of the batched gradient (<= 1: skip)")
    ap.add_argument("--no-coupled", action="store_true", help="skip the coupled two-basin run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 10 device copies of 256 MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        n = 256 * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 10}), flush=True)
        return

    from openqg_amd import workloads as W

    w = W.WORKLOADS[args.workload]()
    qg = w.make_qg(flags=args.flags)
    p = w.initial_p()
    qg.set_state(p=p, pm=p)
    qg.init_state()
    ent, wekp, txis, txin = w.forcing()
    qg.set_forcing(ent, wekp, txis, txin)

    def advance(nsteps, step0):
        if args.eager:   # one launch per kernel, no graph (for counter collection)
            for nt in range(step0 + 1, step0 + nsteps + 1):
                qg.step(w.tdt)
                if (nt - 1) % w.supp_every == 0:
                    qg.suppress()
        else:
            qg.run(w.tdt, nsteps, step0, w.supp_every)

    advance(args.warmup, 0)
    if not args.eager:
        # capture (without running) the graphs the timed run replays, as an
        # FFT library plans before it executes
        qg.plan_run(w.tdt, args.steps, args.warmup, w.supp_every)
    qg.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    advance(args.steps, args.warmup)
    qg.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    days_per_s = args.steps * w.dt / 86400.0 / el

    # per-kernel durations from HIP events on the library's stream: one more
    # step from the timed region's final state, each kernel timed over
    # KERNEL_REPS back-to-back launches on that step's inputs
    kt = qg.kernel_times(w.tdt, KERNEL_REPS)
    dom = max((k for k in kt if kt[k] > 0), key=lambda k: kt[k])
    ach = algorithmic_bytes(w, dom) / (kt[dom] * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(ach, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(ach / HBM_PEAK_GBS, 4),
        "traffic": load_traffic(dom, w.name),
        "kernel_ms": {k: round(v, 5) for k, v in kt.items()},
        # every kernel of the step against the same roof (the longest is the one above)
        "kernel_frac": {k: round(algorithmic_bytes(w, k) / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                        for k, v in kt.items() if v > 0},
        "step_algorithmic_GBs": round(algorithmic_bytes(w, "step") / (ms_per_step * 1e-3) / 1e9, 1),
    }
    progress(f"plain: {days_per_s:.1f} simulated days/s, {ms_per_step:.4f} ms/step")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_seconds)
        progress(f"cpu baseline: {cpu['value']:.2f} simulated days/s")
    qg.close()
    ens = None if args.no_ensemble else ensemble_bench(args.ens_members, args.ens_days)
    if ens:
        progress(f"ensemble: {ens['value']} aggregate simulated days/s")
    adj = None if args.no_adjoint else adjoint_bench(args.adj_days, args.adj_snapshots, args.adj_batch)
    cpl = None if args.no_coupled else coupled_bench(30.0)
    if cpl:
        progress(f"coupled: {cpl['value']} simulated days/s")
    out = {
        "metric": METRIC,
        "value": round(days_per_s, 3),
        "unit": "simulated days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^-3 PV field, double-gyre wind)",
        "config": {"workload": w.name, "nxp": w.nxp, "nyp": w.nyp, "nl": w.nl, "dx_m": w.dx, "dt_s": w.dt,
                   "supp_every": w.supp_every, "y_solver": "thomas" if args.flags & 1 else "wavefront-pcr"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "adjoint": adj,
        "coupled": cpl,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
