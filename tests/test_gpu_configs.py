"""The benchmark configurations at their real sizes on the GPU:

* batched replicas -- the 512-chain mix of the benchmark (N log-uniform in
  [200, 5000], force-ramp pulling), stepped together at T = 0: a seeded
  sample of jobs equals the same jobs run alone on the plain engine, bit
  for bit, and two of them the oracle;
* hydrodynamics at N = 4,096 -- the blocked Cholesky of the 12,288 x 12,288
  RPY matrix in both precisions within 1e-5 covariance error of the FP64
  matrix, and the TEA integrator bit-exact with the oracle for one block;
* serving -- a request mix with an indentation assembly (6,000 beads), a
  force-ramp chain above the LDS step limit and one above the cell-grid
  threshold, a TEA chain and an exact-HI chain, submitted while others run
  against a small HBM budget: every result equals the job alone."""
import numpy as np
import pytest

import bench
import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu


def test_batched_512_chain_mix_equals_plain_at_zero_temperature():
    jobs = bench.build_mix(S, 512, 2024)
    for _, p in jobs:
        p.temperature = 0.0
    numsteps = 100  # stride 10: 110 steps, a possible-pairs rebuild at step 100
    beads = sum(s.n for s, _ in jobs)
    print("case: 512-chain mix, %d beads, T=0, %d steps" % (beads, numsteps + 10), flush=True)
    with S.Engine(0, 64 << 30) as e:
        ids = [e.add_job(p, sysm, numsteps) for sysm, p in jobs]
        L = e.block_steps()
        assert L == 10
        e.run_blocks((numsteps + 10) // L)
        rng = np.random.default_rng(31)
        pick = sorted(rng.choice(len(jobs), size=16, replace=False).tolist())
        batched = {k: e.coords(ids[k]) for k in pick}
        for k in pick:
            st = e.status(ids[k])
            assert st[0] == numsteps + 10 and st[2] == 0
    with S.Engine(0, 8 << 30) as pe:
        for k in pick:
            sysm, p = jobs[k]
            jid = pe.add_job(p, sysm, numsteps)
            pe.run_job(jid)
            np.testing.assert_array_equal(pe.coords(jid), batched[k], err_msg="job %d (N=%d)" % (k, sysm.n))
            pe.remove_job(jid)
    for k in pick[:2]:
        sysm, p = jobs[k]
        op = O.make_params(**{f: (list(getattr(p, f)) if hasattr(getattr(p, f), "__len__") else getattr(p, f))
                              for f, _ in O.OrParams._fields_})
        ref = O.Sim(op, sysm)
        ref.run(numsteps)
        np.testing.assert_array_equal(batched[k], ref.coords(), err_msg="job %d vs oracle" % k)


def _fp64_cov_error(L, D):
    """||L L^T - D||_F / ||D||_F in FP64 (on the GPU through torch)."""
    import torch

    Lt = torch.from_numpy(np.ascontiguousarray(L)).cuda().double()
    Dt = torch.from_numpy(D).cuda()
    e = (torch.linalg.norm(Lt @ Lt.T - Dt) / torch.linalg.norm(Dt)).item()
    del Lt, Dt
    torch.cuda.empty_cache()
    return e


def test_cholesky_order_12288_both_precisions():
    sysm = S.synth_chain(4096, seed=4096)
    D = bench.rpy_matrix(sysm.coords, 3.8)
    A = D.astype(np.float32)
    for mode, name in ((S.CHOL_BF16X3, "split-BF16"), (S.CHOL_F32, "FP32")):
        print("case: Cholesky order 12288, %s" % name, flush=True)
        L, ms = S.chol_factor(A, mode)
        assert np.abs(np.triu(L, 1)).max() == 0.0
        err = _fp64_cov_error(L, D)
        print("%s: %.2f ms, covariance error %.3e" % (name, ms, err), flush=True)
        assert err < 1e-5, (name, err)


def test_tea_4096_beads_bit_exact():
    n = 4096
    sysm = S.synth_chain(n, seed=4097)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    kw = c0_params(pulling=1, fixedEnd=0, pulledEnd=n - 1, fconst=1.43, hi_on=1, hi_a=3.8, hi_epsilon_freq=10,
                   hi_capricious=0, hi_unlisted=1)
    print("case: TEA N=4096, 10 steps", flush=True)
    with S.Engine(0, 4 << 30) as e:
        jid = e.add_job(S.make_params(**kw), sysm, 0)
        e.run_job(jid)
        c, rng = e.coords(jid), e.rng(jid)
        eps, beta = e.hi(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    assert ref.run(0) == 10
    re, rb = ref.hi_tea()
    assert eps == re[0] and beta == rb[0]
    np.testing.assert_array_equal(rng, ref.rng())
    np.testing.assert_array_equal(c, ref.coords())


def _serving_mix():
    out = []
    k = 0

    def ramp(n, steps, seed):
        sysm = S.synth_chain(n, seed=seed) if n > 400 else numpy_chain(n, seed=seed)
        sysm.fixed = np.array([0], np.int32)
        sysm.pulled = np.array([n - 1], np.int32)
        return sysm, c0_params(seed=seed, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02, pullFreq=100,
                               nav=100, outputtiming=100, dcdfreq=100), steps

    out.append(("ramp", *ramp(300, 290, 101)))
    out.append(("ramp-global", *ramp(5500, 190, 102)))
    out.append(("ramp-cells", *ramp(9000, 190, 103)))
    # indentation assembly: 20 chains x 300 beads on a shell, tip above it
    shell = S.synth_shell(20, 300, seed=104, radius=float(np.sqrt(20 * 84.0)))
    c = shell.coords[:, :3]
    top, bot = float(c[:, 2].max()), float(c[:, 2].min())
    out.append(("indentation", shell,
                c0_params(seed=104, indentation=1, indTip=(0.0, 0.0, top + 54.0), indChip=(0.0, 0.0, top + 54.0),
                          indDirection=(0.0, 0.0, -1.0), indSurfR0=(0.0, 0.0, bot - 4.0), indSurfN=(0.0, 0.0, 1.0),
                          indTipR=50.0, indKs=0.07, indDeltaX=0.5, indTipZeta=5000.0, nav=100, outputtiming=100,
                          dcdfreq=100), 290))
    tea = numpy_chain(350, seed=105)
    tea.fixed = np.array([0], np.int32)
    tea.pulled = np.array([349], np.int32)
    out.append(("hi-tea", tea, c0_params(seed=105, pulling=1, fixedEnd=0, pulledEnd=349, fconst=1.43, hi_on=1,
                                         hi_a=1.9, hi_epsilon_freq=10, hi_capricious=0, nav=100, outputtiming=100,
                                         dcdfreq=100), 190))
    ex = S.synth_chain(512, seed=106)
    ex.fixed = np.array([0], np.int32)
    ex.pulled = np.array([511], np.int32)
    out.append(("hi-exact", ex, c0_params(seed=106, pulling=1, fixedEnd=0, pulledEnd=511, fconst=1.43, hi_on=1,
                                          hi_exact=1, hi_a=1.9, hi_chol_freq=10, nav=100, outputtiming=100,
                                          dcdfreq=100), 90))
    out.append(("ramp", *ramp(150, 390, 107)))
    return out


def test_serving_mix_results_equal_alone():
    mix = _serving_mix()
    with S.Server(0, 256 << 20) as srv:  # admission must queue
        tickets = []
        for kind, sysm, kw, steps in mix:
            print("case: serving, submit %s N=%d" % (kind, sysm.n), flush=True)
            tickets.append(srv.submit(S.make_params(**kw), sysm, steps))
            srv.pump(2)
        for _ in range(20000):
            running, queued = srv.pump(5)
            if running == 0 and queued == 0:
                break
        states = [srv.poll(t)[0] for t in tickets]
        assert states == [2] * len(mix), states
        res = [srv.result(t) for t in tickets]
    with S.Engine(0, 4 << 30) as e:
        for (kind, sysm, kw, steps), got in zip(mix, res):
            print("case: serving, %s N=%d alone" % (kind, sysm.n), flush=True)
            jid = e.add_job(S.make_params(**kw), sysm, steps)
            e.run_job(jid)
            np.testing.assert_array_equal(e.coords(jid), got, err_msg=kind)
            e.remove_job(jid)
