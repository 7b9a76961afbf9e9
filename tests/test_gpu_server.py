"""The serving loop: requests of uneven size and protocol (force-ramp,
force-clamp, TEA hydrodynamics, indentation) arriving while others run,
admitted against a small HBM budget, give the same bits as each job alone."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu


def mix():
    jobs = []
    for k, n in enumerate([150, 333, 512, 90, 260, 400]):
        sysm = numpy_chain(n, seed=90 + k)
        sysm.fixed = np.array([0], np.int32)
        sysm.pulled = np.array([n - 1], np.int32)
        kw = c0_params(seed=500 + k, pulling=1, fixedEnd=0, pulledEnd=n - 1, nav=100, outputtiming=100,
                       dcdfreq=100, pullFreq=100)
        if k % 3 == 0:
            kw.update(deltax=0.01 * (k + 1))          # force-ramp
        elif k % 3 == 1:
            kw.update(deltax=0.0, fconst=1.43)         # force-clamp (100 pN)
        else:
            kw.update(deltax=0.0, fconst=1.43, hi_on=1, hi_a=1.9, hi_epsilon_freq=10, hi_capricious=0)
        jobs.append((sysm, kw, 190 + 100 * (k % 3)))
    return jobs


def test_server_results_equal_alone():
    jobs = mix()
    with S.Server(0, 64 << 20) as srv:  # small budget: admission must queue
        tickets = []
        for k, (sysm, kw, steps) in enumerate(jobs):
            tickets.append(srv.submit(S.make_params(**kw), sysm, steps))
            srv.pump(3)  # later requests arrive while earlier ones run
        for _ in range(10000):
            running, queued = srv.pump(5)
            if running == 0 and queued == 0:
                break
        res = [srv.result(t) for t in tickets]
        lat = [srv.poll(t) for t in tickets]
        st = srv.stats()
    assert all(l[0] == 2 for l in lat)
    assert st["finished"] == len(jobs)
    for (sysm, kw, steps), got in zip(jobs, res):
        ref = O.Sim(O.make_params(**kw), sysm)
        ref.run(steps)
        np.testing.assert_array_equal(got, ref.coords())


def test_server_admits_a_larger_job_after_a_smaller_one_retires():
    """Budget below 1 GiB (each arena block gets a chunk of its own size):
    the larger request fits only once the smaller one has retired and its
    chunks are returned, and both finish with the oracle's bits."""
    small, large = numpy_chain(100, seed=95), numpy_chain(400, seed=96)
    ks = c0_params(seed=951)
    kl = c0_params(seed=961)
    bs = S.job_bytes(S.make_params(**ks), small)
    bl = S.job_bytes(S.make_params(**kl), large)
    assert bl > bs
    with S.Server(0, bl + bs // 2) as srv:
        t1 = srv.submit(S.make_params(**ks), small, 190)
        t2 = srv.submit(S.make_params(**kl), large, 190)
        for _ in range(2000):
            running, queued = srv.pump(5)
            if running == 0 and queued == 0:
                break
        assert srv.poll(t1)[0] == 2 and srv.poll(t2)[0] == 2, (srv.poll(t1), srv.poll(t2))
        got = [srv.result(t1), srv.result(t2)]
    for (sysm, kw), c in zip(((small, ks), (large, kl)), got):
        r = O.Sim(O.make_params(**kw), sysm)
        r.run(190)
        np.testing.assert_array_equal(c, r.coords())
