"""The autotuner and the launch shapes it chooses between.

* Tuning runs candidate shapes on the live jobs from a snapshot and
  restores it: afterwards every job's step, coordinates, RNG state,
  temperature accumulator and lists are what they were.
* Every shape gives the same trajectories, bit for bit: unit sizes 512 /
  768 / 1024 rows, shared or one-job units, pair batches 4 / 6 / 8, Verlet
  chunks 4 / 8, row sorts every 10 / 20 / 40 rebuilds, and the global
  kernel's grid, on jobs either side of the LDS step limit, batched and
  alone, and equal to the oracle."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu


def _jobs():
    out = []
    for k, n in enumerate([300, 1100, 2049, 700, 6000]):
        sysm = numpy_chain(n, seed=140 + k) if n <= 400 else S.synth_chain(n, seed=140 + k)
        sysm.fixed = np.array([0], np.int32)
        sysm.pulled = np.array([n - 1], np.int32)
        kw = c0_params(seed=1400 + k, temperature=0.6, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02,
                       pullFreq=100, nav=100, outputtiming=100, dcdfreq=100)
        out.append((sysm, kw))
    return out


def _state(e, ids):
    return [(e.status(j), e.coords(j), e.rng(j), e.temperature(j), e.list(j, 2)[0]) for j in ids]


def test_autotune_leaves_the_jobs_as_they_were():
    jobs = _jobs()
    with S.Engine(0, 4 << 30) as e:
        ids = [e.add_job(S.make_params(**kw), sysm, 790) for sysm, kw in jobs]
        e.run_blocks(13)  # mid-cadence: 130 steps
        before = _state(e, ids)
        print("case: autotune on 5 jobs at step 130", flush=True)
        rep = e.autotune()
        assert rep["trials"] >= 5 and rep["ms_per_step_best"] <= rep["ms_per_step_start"]
        assert e.shape() == rep["shape"]
        after = _state(e, ids)
        for b, a in zip(before, after):
            assert b[0] == a[0]
            for x, y in zip(b[1:], a[1:]):
                np.testing.assert_array_equal(x, y)
        e.run_blocks(67)
        tuned = [e.coords(j) for j in ids]
    with S.Engine(0, 4 << 30) as e:
        ids = [e.add_job(S.make_params(**kw), sysm, 790) for sysm, kw in jobs]
        e.run_blocks(80)
        for j, c in zip(ids, tuned):
            assert e.status(j)[0] == 800
            np.testing.assert_array_equal(e.coords(j), c)


@pytest.mark.parametrize("shape", [
    dict(unit_rows=512, shared_units=0, pair_batch=8, verlet_chunk=4, sort_every=10, max_blocks=256),
    dict(unit_rows=768, shared_units=1, pair_batch=4, verlet_chunk=8, sort_every=40, max_blocks=0),
    dict(unit_rows=128, shared_units=1, pair_batch=6, verlet_chunk=8, sort_every=20, max_blocks=0),
])
def test_every_shape_gives_the_same_bits(shape):
    jobs = _jobs()
    runs = []
    for sh in (None, shape):
        with S.Engine(0, 4 << 30) as e:
            if sh:
                e.set_shape(**sh)
                assert e.shape() == dict(e.shape(), **sh)
            ids = [e.add_job(S.make_params(**kw), sysm, 490) for sysm, kw in jobs]
            print("case: shape %s, batched" % sh, flush=True)
            e.run_blocks(50)
            got = [e.coords(j) for j in ids]
            print("case: shape %s, alone" % sh, flush=True)
            a = e.add_job(S.make_params(**jobs[2][1]), jobs[2][0], 490)
            e.run_job(a)
            got.append(e.coords(a))
            runs.append(got)
    for x, y in zip(*runs):
        np.testing.assert_array_equal(x, y)
    sysm, kw = jobs[0]
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(490)
    np.testing.assert_array_equal(runs[1][0], ref.coords())


def test_shape_values_are_checked():
    with S.Engine(0, 1 << 26) as e:
        for bad in (dict(unit_rows=1000), dict(pair_batch=5), dict(verlet_chunk=6), dict(sort_every=0)):
            with pytest.raises(S.SopError):
                e.set_shape(**bad)
