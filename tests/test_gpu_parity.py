"""GPU parity: the HIP path through the C ABI against the oracle.

Bars: lists, RNG streams, coordinates of whole trajectories and per-bead
energies bit-exact with the oracle's FP32 restatement; forces bit-exact and
within 1e-5 (norm-wise relative) of the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = S.Engine(0, 8 << 30)
    yield e
    e.close()


def lists_equal(a_cnt, a_ent, b_cnt, b_ent):
    if not np.array_equal(a_cnt, b_cnt):
        return False
    for i in range(len(a_cnt)):
        if not np.array_equal(a_ent[: a_cnt[i], i], b_ent[: b_cnt[i], i]):
            return False
    return True


def test_rng_stream_and_gaussians_bit_exact():
    seeds = O.generate_seeds(123, 512)
    u, st = S.rng_probe(seeds, 40, gaussian=False)
    ou, ost = O.ht_stream(seeds[:64], 40)
    np.testing.assert_array_equal(u[:64], ou)
    g, gst = S.rng_probe(seeds, 10, gaussian=True)
    og, _ = O.gaussians(seeds[:64], 10)
    np.testing.assert_array_equal(g[:64], og)


@pytest.mark.parametrize("n,seed", [(64, 1), (257, 2), (1000, 3)])
def test_forces_bit_exact_and_fp64(eng, n, seed):
    sysm = S.synth_chain(n, seed=seed)
    kw = c0_params(pairs_threshold=32.0)
    jid = eng.add_job(S.make_params(**kw), sysm, 0)
    try:
        eng.rebuild(jid, 3)
        ref = O.Sim(O.make_params(**kw), sysm)
        ref.build_possible()
        ref.build_verlet()
        assert lists_equal(*eng.list(jid, 3), *ref.list(3))
        assert lists_equal(*eng.list(jid, 2), *ref.list(2))
        f = eng.forces(jid)
        np.testing.assert_array_equal(f, ref.forces32())
        f64 = ref.forces64()
        err = np.linalg.norm(f[:, :3] - f64) / np.linalg.norm(f64)
        assert err < 1e-5, err
        per, sums = eng.energies(jid)
        np.testing.assert_array_equal(per, ref.energies32())
        np.testing.assert_array_equal(sums, ref.sum_energies()[:, 0])
    finally:
        eng.remove_job(jid)


def test_possible_pairs_with_many_exclusions(eng):
    sysm = numpy_chain(300, seed=9)
    kw = c0_params(pairs_threshold=20.0, pairlist_cutoff=12.0)
    jid = eng.add_job(S.make_params(**kw), sysm, 0)
    try:
        eng.rebuild(jid, 3)
        ref = O.Sim(O.make_params(**kw), sysm)
        ref.build_possible()
        ref.build_verlet()
        assert lists_equal(*eng.list(jid, 3), *ref.list(3))
        assert lists_equal(*eng.list(jid, 2), *ref.list(2))
    finally:
        eng.remove_job(jid)


def test_short_division_equals_ieee_over_all_divisors():
    """The repulsive-pair kernel's short division (hardware reciprocal
    estimate, one Newton step, Markstein's correction) against IEEE a/b:
    every float divisor b in [2^-64, 2^62] (all 1.06e9 bit patterns) with
    64 numerators: 0, 1, sigma^2 = 3.8^2 and 4^2, the ends of the numerator
    range, and log-uniform random ones."""
    rng = np.random.default_rng(77)
    a = [0.0, 1.0, np.float32(3.8) * np.float32(3.8), 16.0, 2.0 ** -60, 2.0 ** 60]
    a += list(np.exp2(rng.uniform(-60, 60, size=58)))
    a = np.array(a, np.float32)
    a = np.clip(a, np.float32(2.0 ** -60), np.float32(2.0 ** 60)) * (a != 0)
    assert S.check_division(0x1F800000, 0x5E800001, a) == 0


def test_possible_pairs_more_than_64_exclusions(eng):
    """Beads with more than 64 exclusions (the list scan takes them 64 at a
    time): bead 5 and bead 250 get ~100 extra native contacts each."""
    sysm = numpy_chain(300, seed=11)
    have = {tuple(sorted(p)) for p in sysm.natives.tolist()}
    extra = [p for p in [(5, j) for j in range(9, 260, 2)] + [(j, 250) for j in range(20, 246, 2)]
             if tuple(sorted(p)) not in have]
    assert sum(1 for p in extra if 5 in p) > 64 and sum(1 for p in extra if 250 in p) > 64
    c = sysm.coords[:, :3].astype(np.float64)
    ex = np.array(extra, np.int32)
    sysm.natives = np.concatenate([sysm.natives, ex])
    sysm.nat_r0 = np.concatenate([sysm.nat_r0, np.linalg.norm(c[ex[:, 0]] - c[ex[:, 1]], axis=1).astype(np.float32)])
    sysm.nat_eh = np.concatenate([sysm.nat_eh, np.ones(len(ex), np.float32)])
    kw = c0_params(pairs_threshold=200.0, pairlist_cutoff=30.0, max_native=256)
    jid = eng.add_job(S.make_params(**kw), sysm, 0)
    try:
        eng.rebuild(jid, 3)
        ref = O.Sim(O.make_params(**kw), sysm)
        ref.build_possible()
        ref.build_verlet()
        assert lists_equal(*eng.list(jid, 3), *ref.list(3))
        assert lists_equal(*eng.list(jid, 2), *ref.list(2))
    finally:
        eng.remove_job(jid)


def _pull_system(n, seed):
    sysm = numpy_chain(n, seed=seed) if n <= 400 else S.synth_chain(n, seed=seed)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    return sysm


@pytest.mark.parametrize("temperature", [0.0, 0.6])
def test_trajectory_bit_exact_plain(eng, temperature):
    n = 400
    sysm = _pull_system(n, 12)
    kw = c0_params(temperature=temperature, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02, pullFreq=100,
                   nav=100, outputtiming=100, dcdfreq=100)
    jid = eng.add_job(S.make_params(**kw), sysm, 990)
    try:
        eng.run_job(jid)
        ref = O.Sim(O.make_params(**kw), sysm)
        assert ref.run(990) == 1000
        assert eng.status(jid)[0] == 1000
        np.testing.assert_array_equal(eng.coords(jid), ref.coords())
        np.testing.assert_array_equal(eng.rng(jid), ref.rng())
        np.testing.assert_array_equal(eng.temperature(jid), ref.temperature())
        np.testing.assert_array_equal(eng.ext(jid), ref.ext())
        assert lists_equal(*eng.list(jid, 2), *ref.list(2))
    finally:
        eng.remove_job(jid)


def test_zero_temperature_1000_steps_vs_fp64(eng):
    n = 1000
    sysm = _pull_system(n, 13)
    kw = c0_params(temperature=0.0, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.05, pullFreq=10)
    jid = eng.add_job(S.make_params(**kw), sysm, 990)
    try:
        eng.run_job(jid)
        x = eng.coords(jid)[:, :3].astype(np.float64)
    finally:
        eng.remove_job(jid)
    x64 = O.Sim(O.make_params(**kw), sysm).run64(990)
    err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    assert err < 1e-5, err


def test_batched_equals_plain_and_alone_equals_mixed():
    """Jobs of uneven sizes and protocols stepped together in one kernel per
    step give the same bits as each job run alone."""
    rng = np.random.default_rng(5)
    sizes = [200, 333, 1000, 64, 517]
    jobs = []
    for k, n in enumerate(sizes):
        sysm = _pull_system(n, 30 + k)
        kw = c0_params(temperature=[0.0, 0.6][k % 2], seed=100 + k, pulling=1, fixedEnd=0, pulledEnd=n - 1,
                       deltax=float(rng.uniform(0.005, 0.05)), k_trans=float(rng.uniform(0.014, 0.1)), pullFreq=100,
                       nav=100, outputtiming=100, dcdfreq=100)
        jobs.append((sysm, kw))
    steps = 490
    alone = []
    with S.Engine(0, 4 << 30) as e1:
        for sysm, kw in jobs:
            jid = e1.add_job(S.make_params(**kw), sysm, steps)
            e1.run_job(jid)
            alone.append(e1.coords(jid))
            e1.remove_job(jid)
    with S.Engine(0, 4 << 30) as e2:
        ids = [e2.add_job(S.make_params(**kw), sysm, steps) for sysm, kw in jobs]
        L = e2.block_steps()
        e2.run_blocks((steps // 10 + 1) * 10 // L)
        for jid, ref in zip(ids, alone):
            assert e2.status(jid)[0] == steps // 10 * 10 + 10
            np.testing.assert_array_equal(e2.coords(jid), ref)


def test_lds_resident_step_equals_global_gathers():
    """Jobs of up to 5120 beads step from coordinates staged in LDS; the
    result is bit-identical to the kernel that gathers from HBM, for sizes
    on either side of the 1024-bead work unit and of the 5120-bead limit,
    batched and alone."""
    sizes = [64, 1000, 1024, 1025, 2049, 5120, 5121]
    jobs = []
    for k, n in enumerate(sizes):
        sysm = _pull_system(n, 60 + k)
        kw = c0_params(temperature=[0.0, 0.6][k % 2], seed=300 + k, pulling=1, fixedEnd=0, pulledEnd=n - 1,
                       deltax=0.02, pullFreq=100, nav=100, outputtiming=100, dcdfreq=100)
        jobs.append((sysm, kw))
    steps = 290
    out = {}
    for lds in (True, False):
        print("case: mixed batch %s, lds=%s" % (sizes, lds), flush=True)
        with S.Engine(0, 4 << 30) as e:
            e.set_lds_step(lds)
            ids = [e.add_job(S.make_params(**kw), sysm, steps) for sysm, kw in jobs]
            L = e.block_steps()
            e.run_blocks((steps // 10 + 1) * 10 // L)
            out[lds] = [e.coords(j) for j in ids]
            for j in ids:
                assert e.status(j)[0] == 300
        print("case: 2049 beads alone, lds=%s" % lds, flush=True)
        with S.Engine(0, 1 << 30) as e:
            e.set_lds_step(lds)
            j = e.add_job(S.make_params(**jobs[4][1]), jobs[4][0], steps)
            e.run_job(j)
            out[(lds, "alone")] = e.coords(j)
    for a, b in zip(out[True], out[False]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(out[(True, "alone")], out[(False, "alone")])
    np.testing.assert_array_equal(out[(True, "alone")], out[True][4])


def test_job_admitted_mid_run_matches_alone():
    n1, n2 = 300, 450
    s1, s2 = _pull_system(n1, 40), _pull_system(n2, 41)
    k1 = c0_params(pulling=1, fixedEnd=0, pulledEnd=n1 - 1, deltax=0.02, pullFreq=100, seed=7)
    k2 = c0_params(pulling=1, fixedEnd=0, pulledEnd=n2 - 1, deltax=0.03, pullFreq=100, seed=8)
    with S.Engine(0, 2 << 30) as e:
        a = e.add_job(S.make_params(**k1), s1, 390)
        e.run_blocks(15)
        b = e.add_job(S.make_params(**k2), s2, 290)
        e.run_blocks(40)
        ca, cb = e.coords(a), e.coords(b)
        assert e.status(a)[0] == 400 and e.status(b)[0] == 300
    ra = O.Sim(O.make_params(**k1), s1); ra.run(390)
    rb = O.Sim(O.make_params(**k2), s2); rb.run(290)
    np.testing.assert_array_equal(ca, ra.coords())
    np.testing.assert_array_equal(cb, rb.coords())


def test_batched_jobs_with_different_list_cadences_match_oracle():
    """Jobs whose Verlet / possible-pairs rebuilds fall on different steps
    (every 10/100, 20/60 and 5/50 steps: blocks of 5 steps in which only
    some jobs rebuild, and row sorts at different steps) stepped together
    from captured blocks, and the first one alone: each trajectory is the
    oracle's, bit for bit."""
    cad = [(300, 10, 100), (400, 20, 60), (250, 5, 50)]
    steps = 400
    jobs = []
    for k, (n, pf, ppf) in enumerate(cad):
        sysm = _pull_system(n, 80 + k)
        kw = c0_params(seed=500 + k, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02, pullFreq=100,
                       nav=100, outputtiming=100, dcdfreq=100, pairs_freq=pf, possiblepairs_freq=ppf)
        jobs.append((sysm, kw))
    with S.Engine(0, 2 << 30) as e:
        ids = [e.add_job(S.make_params(**kw), sysm, steps) for sysm, kw in jobs]
        L = e.block_steps()
        assert L == 5  # strides 10, 20, 5
        ends = [(steps // st + 1) * st for st in (10, 20, 5)]  # the reference's outer loop runs 0..numsteps/stride
        e.run_blocks(max(ends) // L)
        got = [e.coords(j) for j in ids]
        for j, end in zip(ids, ends):
            assert e.status(j)[0] == end
        a = e.add_job(S.make_params(**jobs[0][1]), jobs[0][0], steps)
        e.run_job(a)
        alone = e.coords(a)
    for (sysm, kw), c in zip(jobs, got):
        print("case: oracle, pairs_freq %d possiblepairs_freq %d" % (kw["pairs_freq"], kw["possiblepairs_freq"]),
              flush=True)
        r = O.Sim(O.make_params(**kw), sysm)
        r.run(steps)
        np.testing.assert_array_equal(c, r.coords())
    np.testing.assert_array_equal(alone, got[0])


@pytest.mark.parametrize("n", [300, 1500])
def test_indentation_bit_exact(eng, n):
    """The tip and substrate walls and the tip's Langevin update; the
    reaction on the tip is summed in bead order in chunks of 512 beads
    (1,500: two whole chunks and a partial one)."""
    sysm = numpy_chain(n, seed=50) if n <= 400 else S.synth_chain(n, seed=50)
    c = sysm.coords[:, :3]
    top = float(c[:, 2].max())
    kw = c0_params(temperature=0.6, indentation=1, indTip=(0, 0, top + 53.0), indChip=(0, 0, top + 53.0),
                   indDirection=(0, 0, -1), indSurfR0=(0, 0, float(c[:, 2].min()) - 5.0), indSurfN=(0, 0, 1),
                   indTipR=50.0, indKs=0.07, indDeltaX=0.5, indTipZeta=5000.0, indSurfA=1.0, indSurfB=-2.0,
                   indSurfSigma=2.0, indSurfEl=0.2, nav=50, outputtiming=50, dcdfreq=50)
    jid = eng.add_job(S.make_params(**kw), sysm, 490)
    try:
        eng.run_job(jid)
        ref = O.Sim(O.make_params(**kw), sysm)
        ref.run(490)
        np.testing.assert_array_equal(eng.coords(jid), ref.coords())
        np.testing.assert_array_equal(eng.indentation(jid)[:9], ref.indentation())
    finally:
        eng.remove_job(jid)


def test_pairlist_overflow_is_reported(eng):
    sysm = numpy_chain(200, seed=60)
    kw = c0_params(pairs_threshold=30.0, max_possiblePairs=8)
    jid = eng.add_job(S.make_params(**kw), sysm, 0)
    try:
        eng.rebuild(jid, 1)
        assert eng.status(jid)[2] & 1
    finally:
        eng.remove_job(jid)


def test_verlet_overflow_is_capped_and_reported():
    """A job whose Verlet rows outgrow max_pairs = 16 (a compact 300-bead
    globule has ~100 partners within 20 A) sets overflow bit 2, keeps the
    first 16 entries of every row, as the oracle does, and writes nothing
    past its list; a job batched with it gives the oracle's bits too."""
    dense, other = numpy_chain(300, seed=61), numpy_chain(250, seed=62)
    ka = c0_params(max_pairs=16, seed=611)
    kb = c0_params(seed=622)
    with S.Engine(0, 1 << 30) as e:
        a = e.add_job(S.make_params(**ka), dense, 190)
        b = e.add_job(S.make_params(**kb), other, 190)
        e.run_blocks(200 // e.block_steps())
        assert e.status(a)[2] & 2 and e.status(b)[2] == 0
        cnt, _ = e.list(a, 2)
        assert cnt.max() == 16
        ca, cb = e.coords(a), e.coords(b)
    ra = O.Sim(O.make_params(**ka), dense)
    ra.run(190)
    rb = O.Sim(O.make_params(**kb), other)
    rb.run(190)
    np.testing.assert_array_equal(ca, ra.coords())
    np.testing.assert_array_equal(cb, rb.coords())


@pytest.mark.parametrize("n,lds", [(257, False), (1000, False), (6000, True)])
def test_forces_of_the_global_gather_step_kernel(n, lds):
    """The forces-only pass of the global-gather step kernel (jobs above
    5,120 beads, or the LDS path switched off) equals the oracle's FP32
    forces bit for bit."""
    sysm = S.synth_chain(n, seed=n)
    kw = c0_params(pairs_threshold=32.0)
    with S.Engine(0, 2 << 30) as e:
        e.set_lds_step(lds)
        jid = e.add_job(S.make_params(**kw), sysm, 0)
        e.rebuild(jid, 3)
        f = e.forces(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.build_possible()
    ref.build_verlet()
    np.testing.assert_array_equal(f, ref.forces32())


def test_heating_schedule_bit_exact():
    """The heating schedule (src/Integrators/langevin.cu:68-82: T raised by
    deltaT every tempFreq steps) run on the device in the job updater:
    coordinates, RNG state and temperature accumulator equal the oracle's,
    alone and batched with a job that does not heat."""
    n = 300
    sysm = _pull_system(n, 14)
    kw = c0_params(heating=1, initialT=0.1, deltaT=0.12, tempFreq=60, nav=30, outputtiming=30, dcdfreq=30,
                   pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02, pullFreq=30, seed=4343)
    other = _pull_system(200, 15)
    ko = c0_params(pulling=1, fixedEnd=0, pulledEnd=199, deltax=0.01, pullFreq=100, seed=4344)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(590)
    with S.Engine(0, 1 << 30) as e:
        a = e.add_job(S.make_params(**kw), sysm, 590)
        e.run_job(a)
        np.testing.assert_array_equal(e.coords(a), ref.coords())
        np.testing.assert_array_equal(e.rng(a), ref.rng())
        np.testing.assert_array_equal(e.temperature(a), ref.temperature())
    with S.Engine(0, 1 << 30) as e:
        a = e.add_job(S.make_params(**kw), sysm, 590)
        b = e.add_job(S.make_params(**ko), other, 590)
        L = e.block_steps()
        e.run_blocks(600 // L)
        assert e.status(a)[0] == 600
        np.testing.assert_array_equal(e.coords(a), ref.coords())
