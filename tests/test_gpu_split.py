"""The step with each bead's force terms spread over a workgroup
(split_rows 4, 8, 16): the owners add the terms in the one-lane step's
order, so forces, trajectories and hydrodynamic steps equal the oracle and
the one-bead-per-lane step bit for bit -- on lists long enough for several
chunks, on jobs of a few rows, beside jobs stepped by the global kernel,
batched and alone."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu

SPLITS = [4, 8, 16]


@pytest.mark.parametrize("split", SPLITS)
@pytest.mark.parametrize("n", [70, 1000, 5000])
def test_split_forces_bit_exact(n, split):
    """Forces-only pass after list builds by the spread possible-list and
    Verlet kernels: long Verlet rows (threshold 32 A: several 208-entry
    chunks per row) and a job whose last tile holds 6 rows.  Forces equal
    the oracle's, lists the one-bead-per-lane kernels'."""
    sysm = S.synth_chain(n, seed=900 + n)
    kw = c0_params(pairs_threshold=32.0)
    got = []
    for sh in (split, 0):
        with S.Engine(0, 2 << 30) as e:
            e.set_shape(split_rows=sh)
            jid = e.add_job(S.make_params(**kw), sysm, 0)
            print("case: split lists and forces n=%d R=%d" % (n, sh), flush=True)
            e.rebuild(jid, 3)
            got.append((e.forces(jid), e.list(jid, 1), e.list(jid, 2)))
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.build_possible()
    ref.build_verlet()
    np.testing.assert_array_equal(got[0][0], ref.forces32())
    np.testing.assert_array_equal(got[1][0], ref.forces32())
    for w in (1, 2):
        np.testing.assert_array_equal(got[0][w][0], got[1][w][0])
        np.testing.assert_array_equal(got[0][w][1], got[1][w][1])


@pytest.mark.parametrize("split", SPLITS)
def test_split_trajectory_bit_exact(split):
    """A pulled 300-bead chain at T > 0 (row sorts, list rebuilds, the
    fixed/pulled rules), alone: equal to the oracle."""
    n = 300
    sysm = numpy_chain(n, seed=910)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    kw = c0_params(seed=911, temperature=0.6, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02, pullFreq=50,
                   nav=50, outputtiming=50, dcdfreq=50)
    with S.Engine(0, 1 << 30) as e:
        e.set_shape(split_rows=split)
        jid = e.add_job(S.make_params(**kw), sysm, 390)
        print("case: split trajectory n=300 R=%d" % split, flush=True)
        e.run_job(jid)
        c, rng, T = e.coords(jid), e.rng(jid), e.temperature(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(390)
    np.testing.assert_array_equal(c, ref.coords())
    np.testing.assert_array_equal(rng, ref.rng())
    np.testing.assert_array_equal(T, ref.temperature())


@pytest.mark.parametrize("split", SPLITS)
def test_split_batched_equals_one_lane_step(split):
    """Five jobs either side of the LDS limit (the 6,000-bead job is
    stepped from global memory beside the LDS jobs, spread as well),
    batched, then one job alone: the same bits as the one-bead-per-lane
    step."""
    runs = []
    for sh in (0, split):
        with S.Engine(0, 4 << 30) as e:
            e.set_shape(split_rows=sh, split_global=sh)
            ids = []
            for k, n in enumerate([1000, 130, 2049, 700, 6000]):
                sysm = S.synth_chain(n, seed=920 + k)
                sysm.fixed = np.array([0], np.int32)
                sysm.pulled = np.array([n - 1], np.int32)
                kw = c0_params(seed=930 + k, temperature=0.6, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02,
                               pullFreq=100, nav=100, outputtiming=100, dcdfreq=100)
                ids.append(e.add_job(S.make_params(**kw), sysm, 390))
            print("case: split batched R=%d" % sh, flush=True)
            e.run_blocks(40)
            got = [(e.coords(j), e.rng(j), e.temperature(j)) for j in ids]
            sysm = S.synth_chain(1000, seed=940)
            kw = c0_params(seed=941, temperature=0.6, pairs_threshold=32.0)
            a = e.add_job(S.make_params(**kw), sysm, 190)
            print("case: split alone R=%d" % sh, flush=True)
            e.run_job(a)
            got.append((e.coords(a), e.rng(a), e.temperature(a)))
            runs.append(got)
    for x, y in zip(*runs):
        for u, v in zip(x, y):
            np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("split", [4, 16])
def test_split_tea_bit_exact(split):
    """A TEA hydrodynamic job: the step kernel leaves forces and noise for
    the HI update; equal to the oracle."""
    n = 300
    sysm = numpy_chain(n, seed=950)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    kw = c0_params(hi_on=1, hi_a=1.9, hi_epsilon_freq=10, hi_capricious=0, hi_unlisted=1, pulling=1, fixedEnd=0,
                   pulledEnd=n - 1, deltax=0.02, pullFreq=100, nav=100, outputtiming=100, dcdfreq=100)
    with S.Engine(0, 1 << 30) as e:
        e.set_shape(split_rows=split)
        jid = e.add_job(S.make_params(**kw), sysm, 190)
        print("case: split TEA R=%d" % split, flush=True)
        e.run_job(jid)
        c, rng = e.coords(jid), e.rng(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(190)
    np.testing.assert_array_equal(rng, ref.rng())
    np.testing.assert_array_equal(c, ref.coords())


def test_split_indentation_bit_exact():
    n = 300
    sysm = numpy_chain(n, seed=960)
    c = sysm.coords[:, :3]
    top = float(c[:, 2].max())
    kw = c0_params(temperature=0.6, indentation=1, indTip=(0, 0, top + 53.0), indChip=(0, 0, top + 53.0),
                   indDirection=(0, 0, -1), indSurfR0=(0, 0, float(c[:, 2].min()) - 5.0), indSurfN=(0, 0, 1),
                   indTipR=50.0, indKs=0.07, indDeltaX=0.5, indTipZeta=5000.0, indSurfA=1.0, indSurfB=-2.0,
                   indSurfSigma=2.0, indSurfEl=0.2, nav=50, outputtiming=50, dcdfreq=50)
    with S.Engine(0, 1 << 30) as e:
        e.set_shape(split_rows=8)
        jid = e.add_job(S.make_params(**kw), sysm, 490)
        print("case: split indentation R=8", flush=True)
        e.run_job(jid)
        got, ind = e.coords(jid), e.indentation(jid)[:9]
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(490)
    np.testing.assert_array_equal(got, ref.coords())
    np.testing.assert_array_equal(ind, ref.indentation())


@pytest.mark.parametrize("split", SPLITS)
def test_split_global_forces_bit_exact(split):
    """Jobs above the LDS limit, partners gathered from global memory:
    forces of a 6,000-bead chain equal the oracle's."""
    n = 6000
    sysm = S.synth_chain(n, seed=970)
    kw = c0_params(pairs_threshold=32.0)
    with S.Engine(0, 2 << 30) as e:
        e.set_shape(split_global=split)
        jid = e.add_job(S.make_params(**kw), sysm, 0)
        print("case: split_global forces n=%d R=%d" % (n, split), flush=True)
        e.rebuild(jid, 3)
        f = e.forces(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.build_possible()
    ref.build_verlet()
    np.testing.assert_array_equal(f, ref.forces32())


@pytest.mark.parametrize("split", [4, 16])
def test_split_global_trajectories_equal_one_lane_step(split):
    """A 70,000-bead chain (32-bit list entries) pulled at T > 0, and a
    6,000-bead chain under an indentation tip, each with and without the
    spread step: the same bits."""
    sysm = S.synth_chain(70000, seed=980)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([69999], np.int32)
    kw = c0_params(seed=981, temperature=0.6, pulling=1, fixedEnd=0, pulledEnd=69999, deltax=0.02, pullFreq=50,
                   nav=50, outputtiming=50, dcdfreq=50)
    ind = S.synth_chain(6000, seed=982)
    c = ind.coords[:, :3]
    top = float(c[:, 2].max())
    kwi = c0_params(temperature=0.6, indentation=1, indTip=(0, 0, top + 53.0), indChip=(0, 0, top + 53.0),
                    indDirection=(0, 0, -1), indSurfR0=(0, 0, float(c[:, 2].min()) - 5.0), indSurfN=(0, 0, 1),
                    indTipR=50.0, indKs=0.07, indDeltaX=0.5, indTipZeta=5000.0, indSurfA=1.0, indSurfB=-2.0,
                    indSurfSigma=2.0, indSurfEl=0.2, nav=50, outputtiming=50, dcdfreq=50)
    runs = []
    for sh in (0, split):
        with S.Engine(0, 8 << 30) as e:
            e.set_shape(split_global=sh)
            a = e.add_job(S.make_params(**kw), sysm, 90)
            b = e.add_job(S.make_params(**kwi), ind, 190)
            print("case: split_global trajectories R=%d" % sh, flush=True)
            e.run_job(a)
            e.run_job(b)
            runs.append([e.coords(a), e.rng(a), e.temperature(a), e.coords(b), e.indentation(b)[:9]])
    for x, y in zip(*runs):
        np.testing.assert_array_equal(x, y)


def test_split_rows_values_are_checked():
    with S.Engine(0, 1 << 26) as e:
        for bad in (2, 32, 64, -4):
            with pytest.raises(S.SopError):
                e.set_shape(split_rows=bad)
            with pytest.raises(S.SopError):
                e.set_shape(split_global=bad)
