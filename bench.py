#!/usr/bin/env python3
"""Benchmark: bead-steps/s of the batched engine vs the plain engine.

Workload (one GPU): 512 independent SOP chains with bead counts log-uniform
in [200, 5000] (seeded), each a compact globule with native contacts
|i-j|>2, r<8 A; force-ramp pulling (bead 0 fixed, last bead pulled) with
speeds log-uniform in [0.1, 10] um/s and cantilever springs in [10, 70]
pN/nm; T = 300 K, h = 40 ps, FP32.  All 512 jobs advance together, one step
kernel per step, blocks of 10 steps replayed from a hipGraph.

Printed JSON line: `value` is the batched engine's aggregate bead-steps/s
over the timed region; `plain` is the same mix run one job at a time (a
seeded sample of jobs, each alone); `roofline` is the step kernel's
algorithmic HBM bytes per launch over its HIP-event-timed average duration;
`cpu_baseline` is the oracle (the reference kernel bodies as host loops,
OpenMP over beads) on a bounded sample of the same mix on this host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sop-gpu_amd"))

import numpy as np  # noqa: E402

METRIC = "bead-steps/s on 1 GPU, batched vs plain engine; p50/p99 job latency at fixed load"
HBM_PEAK = 8.0e12
# The list cadence of the mix repeats every PERIOD steps: Verlet rebuilds
# every 10, possible pairs every 100, the Verlet row sort at every 20th
# Verlet rebuild.  Every timed window is a whole number of periods, so it
# holds each kind of block in its steady-state share.
PERIOD = 200


TAU_H_PS = 250.0          # tau_H = 6 pi eta a^3 / kT at a = 3.8 A, water
H_TAU = 0.16              # 40 ps
PN_PER_KCAL_A = 69.48     # 1 kcal/mol/A in pN


def whole_periods(steps):
    return max(1, -(-int(steps) // PERIOD)) * PERIOD


def log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def job_params(S, n, seed, speed_um_s, kappa_pn_nm):
    pull_freq = 1000
    # v = deltax / (pullFreq h tau_H); 1 um/s = 1e-8 A/ps
    deltax = speed_um_s * 1e-8 * pull_freq * H_TAU * TAU_H_PS
    k_trans = kappa_pn_nm * 0.1 / PN_PER_KCAL_A
    return S.make_params(kspring_cov=14.0, R_limit=2.0, a=3.8, el=1.0, pairs_cutoff=20.0, pairlist_cutoff=20.0,
                         pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=H_TAU, zeta=50.0,
                         temperature=0.6, seed=seed, nav=1000, outputtiming=10000, dcdfreq=10000, pulling=1,
                         fixedEnd=0, pulledEnd=n - 1, deltax=float(deltax), k_trans=float(k_trans), pullFreq=pull_freq)


def build_mix(S, njobs, seed):
    rng = np.random.default_rng(seed)
    sizes = np.round(log_uniform(rng, 200, 5000, njobs)).astype(int)
    speeds = log_uniform(rng, 0.1, 10.0, njobs)
    kappas = rng.uniform(10.0, 70.0, njobs)
    jobs = []
    for k in range(njobs):
        n = int(sizes[k])
        sysm = S.synth_chain(n, seed=seed * 100003 + k)
        sysm.fixed = np.array([0], np.int32)
        sysm.pulled = np.array([n - 1], np.int32)
        jobs.append((sysm, job_params(S, n, 1000 + k, speeds[k], kappas[k])))
    return jobs


def serving_mix(S, njobs, seed, steps_lo, steps_hi):
    """Request mix: 70% force-ramp chains (N log-uniform [100, 20000]), 20%
    indentation assemblies (20-100 chains x 300 beads), 10% TEA-HI
    force-clamp chains (N uniform [512, 4096], 100 pN)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(njobs):
        kind = rng.choice(3, p=[0.7, 0.2, 0.1])
        steps = int(log_uniform(rng, steps_lo, steps_hi, 1)[0])
        if kind == 0:
            n = int(log_uniform(rng, 100, 20000, 1)[0])
            sysm = S.synth_chain(n, seed=seed * 7919 + k)
            sysm.fixed = np.array([0], np.int32)
            sysm.pulled = np.array([n - 1], np.int32)
            p = job_params(S, n, 20000 + k, float(log_uniform(rng, 0.1, 10.0, 1)[0]), float(rng.uniform(10, 70)))
        elif kind == 1:
            nch = int(rng.integers(20, 101))
            radius = math.sqrt(nch * 84.0)
            sysm = S.synth_shell(nch, 300, seed=seed * 7919 + k, radius=radius)
            c = sysm.coords[:, :3]
            top, bot = float(c[:, 2].max()), float(c[:, 2].min())
            tip_r = 50.0
            # tip above the assembly moving down at 1 um/s, substrate plane below
            dx = 1.0 * 1e-8 * 1000 * H_TAU * TAU_H_PS
            p = S.make_params(kspring_cov=14.0, R_limit=2.0, pairs_cutoff=20.0, pairlist_cutoff=20.0,
                              pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=H_TAU,
                              temperature=0.6, seed=30000 + k, nav=1000, indentation=1,
                              indTip=(0.0, 0.0, top + tip_r + 4.0), indChip=(0.0, 0.0, top + tip_r + 4.0),
                              indDirection=(0.0, 0.0, -1.0), indSurfR0=(0.0, 0.0, bot - 4.0),
                              indSurfN=(0.0, 0.0, 1.0), indTipR=tip_r, indKs=0.07, indDeltaX=dx,
                              indTipZeta=5000.0)
        else:
            n = int(rng.integers(512, 4097))
            sysm = S.synth_chain(n, seed=seed * 7919 + k)
            sysm.fixed = np.array([0], np.int32)
            sysm.pulled = np.array([n - 1], np.int32)
            p = S.make_params(kspring_cov=14.0, R_limit=2.0, pairs_cutoff=20.0, pairlist_cutoff=20.0,
                              pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=H_TAU,
                              temperature=0.6, seed=40000 + k, nav=1000, pulling=1, fixedEnd=0, pulledEnd=n - 1,
                              fconst=100.0 / PN_PER_KCAL_A, hi_on=1, hi_a=1.9, hi_epsilon_freq=10,
                              hi_capricious=0)
        out.append((("ramp", "indentation", "hi-tea")[kind], sysm, p, steps))
    return out


def serving_tuning(S, mix, count=16):
    """The autotuner on the first `count` requests of the serving mix
    admitted together (a typical set of jobs the server holds at once)."""
    eng = S.Engine(0, 64 << 30)
    for kind, sysm, p, steps in mix[:count]:
        eng.add_job(p, sysm, steps)
    tune = eng.autotune()
    eng.close()
    return tune


def serving_phase(S, args, capacity):
    mix = serving_mix(S, args.serve_jobs, args.seed + 11, args.serve_steps_lo, args.serve_steps_hi)
    tune = serving_tuning(S, mix)
    print(f"# serving autotune: {tune}", file=sys.stderr, flush=True)
    work = np.array([m[1].n * (m[3] // 10 + 1) * 10 for m in mix], dtype=np.float64)
    rate = args.serve_load * capacity / work.mean()  # jobs per second
    rng = np.random.default_rng(args.seed + 12)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(mix)))
    srv = S.Server(0, int(args.serve_budget_gb * (1 << 30)))
    srv.set_shape(tune["shape"])
    srv.set_tuning(tune["shape"]["max_blocks"], not args.no_graph)
    tickets = [None] * len(mix)
    nxt = 0
    t0 = time.perf_counter()
    last_note = 0.0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(mix) and arrivals[nxt] <= now:
            kind, sysm, p, steps = mix[nxt]
            tickets[nxt] = srv.submit(p, sysm, steps)
            nxt += 1
        running, queued = srv.pump(1)
        if now - last_note >= 15.0:
            last_note = now
            print(f"# serving: {now:.0f} s, {nxt} submitted, {running} running, {queued} queued",
                  file=sys.stderr, flush=True)
        if nxt == len(mix) and running == 0 and queued == 0:
            break
        if running == 0 and queued == 0 and nxt < len(mix):
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - (time.perf_counter() - t0))))
    makespan = time.perf_counter() - t0
    lat, failed = [], 0
    for t in tickets:
        st, ts, tb, td = srv.poll(t)
        if st != 2:
            failed += 1
            continue
        lat.append(td - ts)
    stats = srv.stats()
    srv.close()
    lat = np.array(lat)
    kinds = {k: sum(1 for m in mix if m[0] == k) for k in ("ramp", "indentation", "hi-tea")}
    return {
        "jobs": len(mix), "kinds": kinds, "failed": failed,
        "steps_per_job": [args.serve_steps_lo, args.serve_steps_hi],
        "arrival_rate_per_s": rate, "offered_load": args.serve_load,
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "makespan_s": makespan, "bead_steps": stats["bead_steps"],
        "bead_steps_per_s": stats["bead_steps"] / makespan,
        "hbm_budget_gb": args.serve_budget_gb, "peak_admitted_gb": stats["peak_bytes"] / 2 ** 30,
        "tuning": tune,
    }, mix


def plain_serial_measured(S, mix, window):
    """Seconds the same requests take one at a time on the plain engine:
    every request is run alone for `window` steps (whole list periods, from
    its first step) and its measured time per step is scaled to its own
    step count."""
    pe = S.Engine(0, 32 << 30)
    total, measured = 0.0, 0
    for kind, sysm, p, steps in mix:
        run = min(whole_periods(window), (steps // 10 + 1) * 10)
        jid = pe.add_job(p, sysm, run - 10)
        pe.synchronize()
        t0 = time.perf_counter()
        pe.run_job(jid)
        pe.synchronize()
        dt = time.perf_counter() - t0
        pe.remove_job(jid)
        total += dt / run * ((steps // 10 + 1) * 10)
        measured += run
    pe.close()
    return total, measured


def rpy_matrix(coords, a):
    """D_ij / D_ii of the Rotne-Prager-Yamakawa tensor (3n x 3n, FP64):
    r > 2a: (3a/4r)[(1 + 2a^2/3r^2) I + (1 - 2a^2/r^2) rr^T/r^2];
    r <= 2a: (1 - 9r/32a) I + (3r/32a) rr^T/r^2; identity diagonal blocks."""
    x = np.asarray(coords, np.float64)[:, :3]
    n = x.shape[0]
    D = np.zeros((3 * n, 3 * n))
    for b0 in range(0, n, 512):
        b1 = min(n, b0 + 512)
        d = x[None, :, :] - x[b0:b1, None, :]          # (rows, n, 3)
        r = np.sqrt((d * d).sum(-1))
        rs = np.where(r > 0, r, 1.0)
        far = r > 2 * a
        cii = np.where(far, 3 * a / (4 * rs) * (1 + 2 * a * a / (3 * rs * rs)), 1 - 9 * rs / (32 * a))
        crr = np.where(far, 3 * a / (4 * rs) * (1 - 2 * a * a / (rs * rs)), 3 * rs / (32 * a))
        u = d / rs[..., None]
        blk = crr[..., None, None] * u[..., :, None] * u[..., None, :] + cii[..., None, None] * np.eye(3)
        idx = np.arange(b0, b1)
        blk[idx - b0, idx] = np.eye(3)
        D[3 * b0:3 * b1, :] = blk.transpose(0, 2, 1, 3).reshape(3 * (b1 - b0), 3 * n)
    return D


def covariance_error(L, D):
    """||L L^T - D||_F / ||D||_F in FP64 (on the GPU through torch when it
    is there, else numpy)."""
    try:
        import torch
        if torch.cuda.is_available():
            Lt = torch.from_numpy(np.ascontiguousarray(L)).cuda().double()
            Dt = torch.from_numpy(D).cuda()
            e = (torch.linalg.norm(Lt @ Lt.T - Dt) / torch.linalg.norm(Dt)).item()
            del Lt, Dt
            torch.cuda.empty_cache()
            return e
    except ImportError:
        pass
    Ld = L.astype(np.float64)
    return float(np.linalg.norm(Ld @ Ld.T - D) / np.linalg.norm(D))


def hi_phase(S, n=4096, reps=5, steps=50, cov=True):
    """Exact hydrodynamics at N = 4096 (12,288 degrees of freedom): the
    blocked Cholesky of the RPY matrix, split-BF16 vs FP32 MFMA trailing
    updates, timed and checked against the FP64 matrix; then the exact-HI
    integrator (factor every 10 steps, 100 pN force clamp) in both modes,
    and the TEA integrator on the same chain."""
    sysm = S.synth_chain(n, seed=4096)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    D = rpy_matrix(sysm.coords, 3.8)
    D32 = D.astype(np.float32)
    out = {"n_beads": n, "order": 3 * n, "hi_a": 3.8}
    flops = (3 * n) ** 3 / 3.0
    for mode, name in ((S.CHOL_F32, "f32"), (S.CHOL_BF16X3, "bf16x3")):
        L, ms = S.chol_factor(D32, mode, repeats=reps)
        out[name] = {"ms_per_factor": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                     "covariance_error": covariance_error(L, D) if cov else None}
        print(f"# cholesky {name}: {ms:.3f} ms, cov err {out[name]['covariance_error']}",
              file=sys.stderr, flush=True)
        p = S.make_params(kspring_cov=14.0, R_limit=2.0, pairs_cutoff=20.0, pairlist_cutoff=20.0,
                          pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=H_TAU,
                          temperature=0.6, seed=4096, nav=1000, pulling=1, fixedEnd=0, pulledEnd=n - 1,
                          fconst=100.0 / PN_PER_KCAL_A, hi_on=1, hi_exact=1, hi_a=3.8, hi_chol_freq=10)
        eng = S.Engine(0, 16 << 30)
        eng.set_chol_mode(mode)
        jid = eng.add_job(p, sysm, steps + 10)
        tune = eng.autotune()  # launch shapes for this job; it does not advance
        eng.set_tuning(tune["shape"]["max_blocks"], True)
        out[name]["integrator_shape"] = tune["shape"]
        eng.run_blocks(1)  # warm-up block
        eng.synchronize()
        t0 = time.perf_counter()
        eng.run_blocks(steps // eng.block_steps())
        eng.synchronize()
        dt = time.perf_counter() - t0
        done = (steps // eng.block_steps()) * eng.block_steps()
        out[name]["integrator_steps_per_s"] = done / dt
        eng.close()
    out["speedup_bf16x3_over_f32"] = out["f32"]["ms_per_factor"] / out["bf16x3"]["ms_per_factor"]
    # TEA on the same chain and force clamp: epsilon and beta' every 10 steps
    p = S.make_params(kspring_cov=14.0, R_limit=2.0, pairs_cutoff=20.0, pairlist_cutoff=20.0,
                      pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=H_TAU,
                      temperature=0.6, seed=4096, nav=1000, pulling=1, fixedEnd=0, pulledEnd=n - 1,
                      fconst=100.0 / PN_PER_KCAL_A, hi_on=1, hi_exact=0, hi_a=3.8, hi_epsilon_freq=10,
                      hi_capricious=0)
    eng = S.Engine(0, 16 << 30)
    jid = eng.add_job(p, sysm, 10 * steps + 10)
    tune = eng.autotune()
    eng.set_tuning(tune["shape"]["max_blocks"], True)
    eng.run_blocks(1)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.run_blocks(10 * steps // eng.block_steps())
    eng.synchronize()
    dt = time.perf_counter() - t0
    eps, beta = eng.hi(jid)
    out["tea"] = {"integrator_steps_per_s": (10 * steps // eng.block_steps()) * eng.block_steps() / dt,
                  "epsilon": eps, "beta": beta, "shape": tune["shape"]}
    eng.close()
    return out


def capsid_phase(S, steps=2000):
    """Continuous indentation of a 100,000-bead capsid-like shell (200 chains
    x 500 beads on a sphere of R = 140 A), tip R = 50 A moving at 1 um/s,
    Verlet/possible rebuilds every 10/100 steps: stepping rate with the
    cell-binned possible-pairs build, and one possible-pairs rebuild with
    cells vs the all-pairs scan."""
    sysm = S.synth_shell(200, 500, seed=2026, radius=140.0)
    c = sysm.coords[:, :3]
    top, bot = float(c[:, 2].max()), float(c[:, 2].min())
    tip_r = 50.0
    dx = 1.0 * 1e-8 * 1000 * H_TAU * TAU_H_PS  # 1 um/s, tip moved every nav = 1000 steps
    p = S.make_params(kspring_cov=14.0, R_limit=2.0, pairs_cutoff=20.0, pairlist_cutoff=20.0, pairs_threshold=32.0,
                      pairs_freq=10, possiblepairs_freq=100, max_possiblePairs=1024, timestep=H_TAU,
                      temperature=0.6, seed=2026, nav=1000, indentation=1,
                      indTip=(0.0, 0.0, top + tip_r + 4.0), indChip=(0.0, 0.0, top + tip_r + 4.0),
                      indDirection=(0.0, 0.0, -1.0), indSurfR0=(0.0, 0.0, bot - 4.0), indSurfN=(0.0, 0.0, 1.0),
                      indTipR=tip_r, indKs=0.07, indDeltaX=dx, indTipZeta=5000.0, indTipA=0.0, indTipB=1.0)
    eng = S.Engine(0, 32 << 30)
    jid = eng.add_job(p, sysm, steps + 100)
    tune = eng.autotune()  # the job does not advance
    eng.set_tuning(tune["shape"]["max_blocks"], True)
    eng.run_blocks(10)  # warm-up: 100 steps incl. one possible-pairs rebuild
    eng.synchronize()
    L = eng.block_steps()
    nb = steps // L
    t0 = time.perf_counter()
    eng.run_blocks(nb)
    eng.synchronize()
    dt = time.perf_counter() - t0
    out = {"beads": sysm.n, "natives": int(len(sysm.nat_r0)), "steps": nb * L,
           "bead_steps_per_s": sysm.n * nb * L / dt, "ms_per_step": dt / (nb * L) * 1e3, "tuning": tune}
    for cells, name in ((True, "possible_rebuild_ms_cells"), (False, "possible_rebuild_ms_all_pairs")):
        eng.set_cells(enabled=cells)
        eng.rebuild(jid, 1)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.rebuild(jid, 1)
        eng.synchronize()
        out[name] = (time.perf_counter() - t0) * 1e3
    st = eng.status(jid)
    out["overflow"] = st[2]
    eng.close()
    return out


def bytes_per_bead_step(avg):
    """Algorithmic HBM bytes per bead-step of the fused step kernel:
    coordinates read+write 32, RNG state read+write 32, three list counts 12,
    covalent 8/entry (index + r0), native 12/entry (index + r0^2, -12eh/r0^2),
    Verlet 4/entry, temperature accumulator read+write 8, extForce 16."""
    n_b, n_n, n_p = avg[0], avg[1], avg[2]
    return 32 + 32 + 12 + 8 * n_b + 12 * n_n + 4 * n_p + 8 + 16


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def _oracle_params(O, p):
    return O.make_params(**{f: list(getattr(p, f)) if hasattr(getattr(p, f), "__len__") else getattr(p, f)
                            for f, _ in O.OrParams._fields_})


def cpu_rate(O, sysm, p, target_s, threads):
    """bead-steps/s of the oracle (the reference kernel bodies as host
    loops, OpenMP over beads) stepping one job: a first 200-step list
    period sizes a window of whole periods of about target_s seconds."""
    sim = O.Sim(_oracle_params(O, p), sysm)
    t = sim.bench(PERIOD, threads)
    periods = max(1, int(target_s / max(t, 1e-6)))
    t = sim.bench(periods * PERIOD, threads)
    return sysm.n * periods * PERIOD / t, periods * PERIOD


def cpu_baseline(jobs, c0, target_s, threads):
    """The oracle on this host (test infrastructure), all threads and one
    thread: four jobs of the 512-chain mix run one after another, and the
    single N = 1,000 chain."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rng = np.random.default_rng(7)
    pick = rng.choice(len(jobs), size=min(4, len(jobs)), replace=False)
    out = {}
    for nt in (threads, 1):
        bs, tt = 0.0, 0.0
        for k in pick:
            sysm, p = jobs[k]
            r, steps = cpu_rate(O, sysm, p, target_s / len(pick) / 2, nt)
            bs += sysm.n * steps
            tt += sysm.n * steps / r
        out[nt] = bs / tt
    c0_all, c0_steps = cpu_rate(O, c0[0], c0[1], target_s / 4, threads)
    c0_one, _ = cpu_rate(O, c0[0], c0[1], target_s / 4, 1)
    nproc, model = cpu_info()
    sample = (f"{len(pick)} jobs of the mix (N={','.join(str(jobs[k][0].n) for k in pick)}) one after another, "
              f"whole 200-step list periods (~{target_s / 2:.0f} s per thread count, rebuilds included)")
    return {
        "value": out[threads], "unit": "bead-steps/s", "cores": threads, "kind": "port", "sample": sample,
        "one_thread_value": out[1], "nproc": nproc, "cpu_model": model,
        "c0": {"n_beads": c0[0].n, "value_all_threads": c0_all, "value_one_thread": c0_one,
               "sample_steps": c0_steps, "note": "10^5-step run timed on a sample of whole periods"},
    }


def c0_system(S):
    """One chain of N = 1,000 beads (seeded compact globule, natives
    |i-j| > 2, r < 8 A), bead 0 fixed and bead 999 pulled at 1 um/s by a
    35 pN/nm spring, T = 300 K, h = 40 ps."""
    n = 1000
    sysm = S.synth_chain(n, seed=1000)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    return sysm, job_params(S, n, 1000, 1.0, 35.0)


def c0_phase(S, c0, steps):
    """The N = 1,000 chain for `steps` steps alone on the GPU: the plain
    engine (one launch per kernel) and the batched engine with one job
    (graph replay)."""
    sysm, p = c0
    out = {"n_beads": sysm.n, "steps": steps}
    for name, graph in (("plain", None), ("batched_graph", True)):
        eng = S.Engine(0, 4 << 30)
        jid = eng.add_job(p, sysm, steps - 10)
        tune = eng.autotune()  # launch shapes for this one job; it does not advance
        out["tuning"] = tune
        if graph:
            eng.set_tuning(tune["shape"]["max_blocks"], True)
        eng.synchronize()
        t0 = time.perf_counter()
        if graph:
            eng.run_blocks(steps // eng.block_steps())
        else:
            eng.run_job(jid)
        eng.synchronize()
        dt = time.perf_counter() - t0
        assert eng.status(jid)[0] == steps
        eng.close()
        out[name] = {"bead_steps_per_s": sysm.n * steps / dt, "us_per_step": dt / steps * 1e6}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000,
                    help="timed steps, rounded up to whole 200-step list periods")
    ap.add_argument("--warmup", type=int, default=200, help="untimed steps, rounded up to whole list periods")
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain-jobs", type=int, default=12)
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--c0-steps", type=int, default=100000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape for the batched mix instead of the autotuner's, e.g. "
                         "unit_rows=1024,shared_units=1,pair_batch=6,verlet_chunk=8,sort_every=10 "
                         "(profiler runs: the tuner's device snapshots are many small copies)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-hi", action="store_true")
    ap.add_argument("--no-capsid", action="store_true")
    ap.add_argument("--no-c0", action="store_true")
    ap.add_argument("--c0-only", action="store_true",
                    help="run only the N = 1,000 single-chain phase (profiling)")
    ap.add_argument("--capsid-only", action="store_true",
                    help="run only the 100,000-bead capsid indentation phase (profiling)")
    ap.add_argument("--hi-only", action="store_true",
                    help="run only the exact-HI phase (Cholesky timing and integrator, no covariance check) "
                         "and print its JSON; for profiling the factorisation")
    ap.add_argument("--serve-jobs", type=int, default=40)
    ap.add_argument("--serve-steps-lo", type=int, default=10000)
    ap.add_argument("--serve-steps-hi", type=int, default=100000)
    ap.add_argument("--serve-load", type=float, default=0.6)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-plain-window", type=int, default=1000,
                    help="steps each serving request is timed alone on the plain engine")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build"}))
        sys.exit(2)

    import sopgpu as S

    if args.hi_only:
        print(json.dumps({"hi_cholesky": hi_phase(S, cov=False)}))
        return
    if args.c0_only:
        print(json.dumps({"c0_single_chain": c0_phase(S, c0_system(S), args.c0_steps)}))
        return
    if args.capsid_only:
        print(json.dumps({"capsid_indentation": capsid_phase(S)}))
        return

    t_build = time.time()
    jobs = build_mix(S, args.jobs, args.seed)
    t_build = time.time() - t_build
    beads = sum(j[0].n for j in jobs)
    print(f"# mix: {len(jobs)} jobs, {beads} beads, built in {t_build:.1f} s", file=sys.stderr, flush=True)

    copy_gbps = S.copy_bandwidth(1 << 30, 20)
    print(f"# streaming copy: {copy_gbps:.0f} GB/s", file=sys.stderr, flush=True)
    eng = S.Engine(0, 96 << 30)
    ids = [eng.add_job(p, sysm, 1000000) for sysm, p in jobs]
    L = eng.block_steps()
    assert PERIOD % L == 0
    if args.shape:  # a fixed launch shape instead of the autotuner's search
        eng.set_shape(**{k: int(v) for k, v in (kv.split("=") for kv in args.shape.split(","))})
        tune = {"shape": eng.shape(), "fixed": True}
    else:
        tune = eng.autotune()  # the jobs do not advance
    tuned = tune["shape"]["max_blocks"]
    eng.set_tuning(tuned, not args.no_graph)
    print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    # warm-up: whole list periods, so every block variant (Verlet only,
    # with possible pairs, with the row sort) is captured before timing
    warm = whole_periods(args.warmup)
    eng.run_blocks(warm // L)
    eng.synchronize()
    steps = whole_periods(args.steps)
    nblocks = steps // L
    print(f"# L={L}; warm-up {warm} steps; timing {steps} steps", file=sys.stderr, flush=True)

    # timed region: graph replays over whole list periods
    eng.synchronize()
    t0 = time.perf_counter()
    eng.run_blocks(nblocks)
    eng.synchronize()
    dt = time.perf_counter() - t0
    value = beads * steps / dt

    # roofline: whole periods again, eager, HIP events on the engine's
    # stream around every block's ten step-kernel launches (nothing else is
    # launched between them)
    eng.profile(True)
    prof_steps = whole_periods(min(steps, 2000))
    eng.synchronize()
    t_eager = time.perf_counter()
    eng.run_blocks(prof_steps // L)
    eng.synchronize()
    t_eager = time.perf_counter() - t_eager
    launches, ms = eng.profile_read()
    eng.profile(False)
    avg_ms = ms / max(launches, 1)
    avg_lists, nbeads = eng.list_stats()
    bpb = bytes_per_bead_step(avg_lists)
    achieved = bpb * nbeads / (avg_ms * 1e-3)
    traffic = None
    kernel = "sop::k_step_lds"  # every job of the mix is at most 5120 beads: the LDS step kernel
    tfile = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            if tj.get("kernel") == kernel and tj.get("bytes_per_launch_per_bead") is not None:
                traffic = tj["bytes_per_launch_per_bead"] * nbeads
        except (ValueError, OSError):
            traffic = None
    eng.close()

    # plain engine: a seeded sample of the same jobs, each run alone over
    # whole list periods from its first step
    rng = np.random.default_rng(args.seed + 1)
    pick = rng.choice(len(jobs), size=min(args.plain_jobs, len(jobs)), replace=False)
    plain_steps = whole_periods(args.plain_steps)
    pe = S.Engine(0, 32 << 30)
    plain_bs, plain_t = 0.0, 0.0
    for k in pick:
        sysm, p = jobs[k]
        jid = pe.add_job(p, sysm, plain_steps - 10)
        pe.synchronize()
        t0 = time.perf_counter()
        pe.run_job(jid)
        pe.synchronize()
        plain_t += time.perf_counter() - t0
        plain_bs += sysm.n * plain_steps
        pe.remove_job(jid)
    pe.close()
    plain_value = plain_bs / plain_t

    serving = None
    if not args.no_serve:
        print("# serving phase", file=sys.stderr, flush=True)
        serving, smix = serving_phase(S, args, value)
        serial_s, measured = plain_serial_measured(S, smix, args.serve_plain_window)
        serving["plain_serial_s"] = serial_s
        serving["plain_serial_method"] = (f"each request alone on the plain engine for min(its steps, "
                                          f"{whole_periods(args.serve_plain_window)}) steps, time per step "
                                          f"scaled to its step count ({measured} steps timed)")
        serving["speedup_vs_plain_serial"] = serial_s / serving["makespan_s"]
        print(f"# serving: p50 {serving['p50_latency_s']:.3f} s p99 {serving['p99_latency_s']:.3f} s, "
              f"{serving['speedup_vs_plain_serial']:.2f}x serial plain", file=sys.stderr, flush=True)

    c0 = c0_system(S)
    c0_gpu = None
    if not args.no_c0:
        print("# C0 phase", file=sys.stderr, flush=True)
        c0_gpu = c0_phase(S, c0, args.c0_steps)

    capsid = None
    if not args.no_capsid:
        print("# capsid indentation phase", file=sys.stderr, flush=True)
        capsid = capsid_phase(S)

    hi = None
    if not args.no_hi:
        print("# exact-HI phase", file=sys.stderr, flush=True)
        hi = hi_phase(S)

    cpu = None
    if not args.no_cpu:
        print("# CPU baseline", file=sys.stderr, flush=True)
        threads = os.cpu_count() or 1
        cpu = cpu_baseline(jobs, c0, args.cpu_seconds, threads)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": dt / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched replicas: 512 SOP chains, N log-uniform [200,5000], force-ramp 0.1-10 um/s, "
                        "10-70 pN/nm, T=300K, h=40ps, Verlet/possible rebuild 10/100 steps",
            "jobs": len(jobs), "beads": beads, "block_steps": L, "max_blocks": tuned,
            "graph": not args.no_graph, "steps_requested": args.steps, "warmup_requested": args.warmup,
            "tuning": tune,
            "window": f"whole {PERIOD}-step list periods (steps and warm-up rounded up)",
        },
        "plain": {"value": plain_value, "jobs_sampled": int(len(pick)), "steps_per_job": plain_steps,
                  "batched_over_plain": value / plain_value},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": "profiles/step_kernel_traffic.json (rocprofv3 --pmc FETCH_SIZE and "
                                       "WRITE_SIZE of this bench, scaled by the copy kernel's counters; per bead "
                                       "x beads of this run)",
                     "kernel": kernel, "avg_launch_ms": avg_ms, "launches": launches,
                     "timing": f"HIP events around each block's step launches, {prof_steps} eager steps",
                     "eager_ms_per_step": t_eager / prof_steps * 1e3,
                     "bytes_per_bead_step": bpb,
                     "copy_kernel_gbps": copy_gbps, "frac_of_copy": achieved / (copy_gbps * 1e9),
                     "avg_lists": {"covalent": avg_lists[0], "native": avg_lists[1], "verlet": avg_lists[2],
                                   "possible": avg_lists[3]}},
        "serving": serving,
        "c0_single_chain": c0_gpu,
        "capsid_indentation": capsid,
        "hi_cholesky": hi,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
