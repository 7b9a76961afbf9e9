"""Finite-temperature ensemble observables of force-ramp pulling: the HIP
engine's replicas and the oracle's replicas, run from disjoint seed sets,
agree on the mean peak pulling force and the mean fraction of native
contacts lost within two standard errors of the difference.

The two ensembles share no trajectory, so this checks the statistics of the
dynamics (noise amplitude, force field, pulling protocol) rather than bit
patterns.  Both observables are computed here from the coordinates at the
same steps: the pulling force as k_trans (cantilever - pulled bead), with the
cantilever at x0 + deltax floor(step/pullFreq) along the fixed-to-pulled
direction (pulling.cu:152-165,198-214).  Seeds are fixed, so the outcome is
deterministic; as for any two-sample test at two standard errors, about one
seed choice in twenty would fail by chance alone.  The first replicas are
also run on the oracle with the GPU's own seeds, where the observables must
be identical."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu

N_BEADS = 100
REPLICAS_GPU = 96
REPLICAS_ORACLE = 48
SAME_SEED_CHECKS = 4
STEPS = 6000
NAV = 100
DELTAX = 1.0
K_TRANS = 0.05


def _system():
    sysm = numpy_chain(N_BEADS, seed=77)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([N_BEADS - 1], np.int32)
    return sysm


def _params(seed):
    # 1 A of cantilever travel per 100 steps, k_trans 0.05 kcal/mol/A^2
    return c0_params(temperature=0.6, seed=seed, pulling=1, fixedEnd=0, pulledEnd=N_BEADS - 1, deltax=DELTAX,
                     k_trans=K_TRANS, pullFreq=NAV, nav=NAV, outputtiming=NAV, dcdfreq=NAV)


def _native_fraction_lost(coords, sysm, r_limit=2.0):
    """1 - Q with the reference's folded-contact rule |r - r0| <= R_limit
    (output_manager.cu), over the .top native contacts."""
    c = coords[:, :3].astype(np.float64)
    i, j = sysm.natives[:, 0], sysm.natives[:, 1]
    r = np.linalg.norm(c[i] - c[j], axis=1)
    kept = np.abs(sysm.nat_r0.astype(np.float64) - r) <= r_limit
    return 1.0 - kept.mean()


def _pull_force(coords, x0, step):
    """|k_trans (cantilever - pulled bead)| at `step` (a multiple of pullFreq)."""
    c0 = x0[:, :3].astype(np.float64)
    v = c0[N_BEADS - 1] - c0[0]
    v /= np.linalg.norm(v)
    cant = c0[N_BEADS - 1] + DELTAX * (step // NAV) * v
    return K_TRANS * np.linalg.norm(cant - coords[N_BEADS - 1, :3].astype(np.float64))


def _stats(x):
    x = np.asarray(x, np.float64)
    return x.mean(), x.std(ddof=1) / np.sqrt(len(x))


def _oracle_replica(sysm, seed):
    sim = O.Sim(O.make_params(**_params(seed)), sysm)
    peak = 0.0
    for b in range(1, STEPS // NAV + 1):
        sim.run(sim.step() + NAV - 1)
        assert sim.step() == b * NAV
        peak = max(peak, _pull_force(sim.coords(), sysm.coords, b * NAV))
    return peak, _native_fraction_lost(sim.coords(), sysm)


def test_pulling_ensemble_matches_oracle():
    sysm = _system()
    print("case: GPU ensemble, %d replicas x %d beads x %d steps" % (REPLICAS_GPU, N_BEADS, STEPS), flush=True)
    gpu_peak, gpu_lost = [], []
    with S.Engine(0, 2 << 30) as e:
        ids = [e.add_job(S.make_params(**_params(1000 + r)), sysm, STEPS) for r in range(REPLICAS_GPU)]
        L = e.block_steps()
        assert NAV % L == 0
        peaks = np.zeros(REPLICAS_GPU)
        for b in range(1, STEPS // NAV + 1):
            e.run_blocks(NAV // L)
            for k, jid in enumerate(ids):
                assert e.status(jid)[0] == b * NAV
                peaks[k] = max(peaks[k], _pull_force(e.coords(jid), sysm.coords, b * NAV))
        for k, jid in enumerate(ids):
            gpu_peak.append(peaks[k])
            gpu_lost.append(_native_fraction_lost(e.coords(jid), sysm))
    print("case: oracle ensemble", flush=True)
    or_peak, or_lost = [], []
    for r in range(REPLICAS_ORACLE):
        peak, lost = _oracle_replica(sysm, 5000 + r)
        or_peak.append(peak)
        or_lost.append(lost)
    for r in range(SAME_SEED_CHECKS):
        assert _oracle_replica(sysm, 1000 + r) == (gpu_peak[r], gpu_lost[r])
    for name, a, b in (("peak force", gpu_peak, or_peak), ("natives lost", gpu_lost, or_lost)):
        ma, sa = _stats(a)
        mb, sb = _stats(b)
        se = np.hypot(sa, sb)
        print("%s: GPU %.4f +- %.4f, oracle %.4f +- %.4f" % (name, ma, sa, mb, sb), flush=True)
        assert sa > 0 and sb > 0  # the ensembles do fluctuate
        assert abs(ma - mb) <= 2.0 * se, (name, ma, mb, se)
