"""GPU hydrodynamics against the oracle: TEA bit-exact (coordinates, RNG,
epsilon, beta'), the exact path bit-exact at zero temperature (D.F), the
blocked Cholesky within the stated covariance error of an FP64 factor, and
the split-BF16 and FP32 paths against each other."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params, numpy_chain

pytestmark = pytest.mark.gpu


def hi_kw(**kw):
    d = c0_params(hi_on=1, hi_a=1.9, hi_epsilon_freq=10, hi_capricious=0)
    d.update(kw)
    return d


@pytest.mark.parametrize("unlisted", [1, 0])
def test_tea_bit_exact(unlisted):
    n = 300
    sysm = numpy_chain(n, seed=70)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    kw = hi_kw(hi_unlisted=unlisted, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.02, pullFreq=100, nav=100,
               outputtiming=100, dcdfreq=100)
    with S.Engine(0, 1 << 30) as e:
        jid = e.add_job(S.make_params(**kw), sysm, 190)
        e.run_job(jid)
        c, rng = e.coords(jid), e.rng(jid)
        eps, beta = e.hi(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(190)
    re, rb = ref.hi_tea()
    assert eps == re[0] and beta == rb[0]
    np.testing.assert_array_equal(rng, ref.rng())
    np.testing.assert_array_equal(c, ref.coords())


@pytest.mark.parametrize("mode", [S.CHOL_BF16X3, S.CHOL_F32])
def test_exact_zero_temperature_bit_exact(mode):
    n = 100
    sysm = numpy_chain(n, seed=71)
    kw = hi_kw(hi_exact=1, temperature=0.0)
    with S.Engine(0, 1 << 30) as e:
        e.set_chol_mode(mode)
        jid = e.add_job(S.make_params(**kw), sysm, 40)
        e.run_job(jid)
        c = e.coords(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(40)
    np.testing.assert_array_equal(c, ref.coords())


@pytest.mark.parametrize("n,mode", [(128, S.CHOL_BF16X3), (128, S.CHOL_F32), (700, S.CHOL_BF16X3),
                                    (700, S.CHOL_F32)])
def test_engine_factor_covariance(n, mode):
    """The factor the engine keeps for an exact-HI job: refactorised only at
    step 0 (hi_chol_freq beyond the run), so it is the factor of the initial
    structure's RPY matrix, whose FP64 factor it must match."""
    sysm = S.synth_chain(n, seed=72)
    kw = hi_kw(hi_exact=1, hi_a=1.9, temperature=0.6, hi_chol_freq=1000)
    with S.Engine(0, 4 << 30) as e:
        e.set_chol_mode(mode)
        jid = e.add_job(S.make_params(**kw), sysm, 0)
        e.run_job(jid)  # 10 steps, the factor is of step 0
        L = e.hi_factor(jid).astype(np.float64)
    assert np.abs(np.triu(L, 1)).max() == 0.0
    D = O.rpy_matrix(sysm.coords, 1.9)
    L64 = np.linalg.cholesky(D)
    cov = np.linalg.norm(L @ L.T - D) / np.linalg.norm(D)
    fac = np.linalg.norm(L - L64) / np.linalg.norm(L64)
    print(f"n={n} mode={mode}: covariance error {cov:.3e}, factor error {fac:.3e}")
    assert cov < 1e-5, cov
    assert fac < 1e-4, fac


@pytest.mark.parametrize("n", [384, 1536, 3 * 1000])
def test_standalone_cholesky_covariance_error(n):
    sysm = S.synth_chain(n // 3, seed=73)
    D = O.rpy_matrix(sysm.coords, 3.8)
    L64 = np.linalg.cholesky(D)
    nd = np.linalg.norm(D)
    for mode, tol in ((S.CHOL_BF16X3, 1e-5), (S.CHOL_F32, 1e-5)):
        L, ms = S.chol_factor(D.astype(np.float32), mode)
        L = L.astype(np.float64)
        cov_err = np.linalg.norm(L @ L.T - D) / nd
        fac_err = np.linalg.norm(L - L64) / np.linalg.norm(L64)
        print(f"n={n} mode={mode} covariance error {cov_err:.3e} factor error {fac_err:.3e} {ms:.3f} ms")
        assert cov_err < tol, (mode, cov_err)


@pytest.mark.parametrize("faithful", [0, 1])
def test_exact_noise_statistics_vs_oracle(faithful):
    """Finite temperature: the exact path's displacement statistics agree
    with the oracle's over a short run (same seeds, different rounding)."""
    n = 40
    sysm = numpy_chain(n, seed=74)
    kw = hi_kw(hi_exact=1, temperature=0.6, hi_faithful_blocks=faithful)
    with S.Engine(0, 1 << 30) as e:
        jid = e.add_job(S.make_params(**kw), sysm, 0)
        e.run_job(jid)
        c = e.coords(jid)
    ref = O.Sim(O.make_params(**kw), sysm)
    ref.run(0)
    # one block of 10 steps from identical seeds: trajectories agree to FP32
    # rounding of the factor
    np.testing.assert_allclose(c, ref.coords(), rtol=0, atol=5e-4)


@pytest.mark.parametrize("faithful", [1, 0])
@pytest.mark.parametrize("mode", [S.CHOL_BF16X3, S.CHOL_F32])
def test_exact_finite_temperature_step_vs_oracle(faithful, mode):
    """One exact-HI step at T = 300 K from a known state, 200 beads (order
    600): the displacement is (h/zeta)(D F + B rho g) with B the factor read
    block-symmetrically (the reference's default, bdhitea_kernel.cu:322-337)
    or the factor itself; the GPU's equals the oracle's and the FP64
    statement to the factor's rounding."""
    n = 200
    sysm = numpy_chain(n, seed=75)
    kw = hi_kw(hi_exact=1, temperature=0.6, hi_faithful_blocks=faithful, nav=1, outputtiming=1, dcdfreq=1,
               pairs_freq=1, possiblepairs_freq=1, hi_epsilon_freq=1)
    x0 = sysm.coords[:, :3].astype(np.float64)
    with S.Engine(0, 1 << 30) as e:
        e.set_chol_mode(mode)
        jid = e.add_job(S.make_params(**kw), sysm, 0)
        e.run_job(jid)
        assert e.status(jid)[0] == 1
        dx_gpu = e.coords(jid)[:, :3].astype(np.float64) - x0
    ref = O.Sim(O.make_params(**kw), sysm)
    g, _ = O.gaussians(ref.rng(), 1)
    f = O.Sim(O.make_params(**kw), sysm)
    f.build_possible()
    f.build_verlet()
    F = f.forces64()
    assert ref.run(0) == 1
    dx_ref = ref.coords()[:, :3].astype(np.float64) - x0
    D = O.rpy_matrix(sysm.coords, kw["hi_a"])
    L = np.linalg.cholesky(D)
    if faithful:
        B = np.zeros_like(L)
        for i in range(n):
            for j in range(i + 1):
                blk = L[3 * i:3 * i + 3, 3 * j:3 * j + 3]
                B[3 * i:3 * i + 3, 3 * j:3 * j + 3] = np.tril(blk) + np.tril(blk, -1).T
    else:
        B = L
    rho = np.sqrt(2 * 0.6 * kw["zeta"] / kw["timestep"])
    expect = kw["timestep"] / kw["zeta"] * (D @ F.ravel() + B @ (rho * g[:, 0, :].astype(np.float64)).ravel())
    expect = expect.reshape(n, 3)
    ulp = float(np.spacing(np.float32(np.abs(x0).max() + 1.0)))
    for got in (dx_gpu, dx_ref):
        np.testing.assert_allclose(got, expect, rtol=5e-5, atol=4 * ulp)
    np.testing.assert_allclose(dx_gpu, dx_ref, rtol=5e-5, atol=4 * ulp)


@pytest.mark.parametrize("mode", [S.CHOL_BF16X3, S.CHOL_F32])
def test_graph_replay_with_exact_hi_job_equals_direct_launches(mode):
    """A batch holding an exact-HI job is replayed from captured graphs keyed
    by the block's launch sequence: with refactorisation every 15 steps and
    10-step blocks the sequence cycles through three keys (refactor at step
    0, at step 5, none).  Every job's coordinates after 150 steps at T = 300
    K equal those of the same batch launched kernel by kernel."""
    n_exact, n_ramp, n_tea = 300, 500, 200
    jobs = []
    s = numpy_chain(n_exact, seed=81)
    jobs.append((S.make_params(**hi_kw(hi_exact=1, hi_chol_freq=15, temperature=0.6, seed=5)), s))
    s = numpy_chain(n_ramp, seed=82)
    s.fixed = np.array([0], np.int32)
    s.pulled = np.array([n_ramp - 1], np.int32)
    jobs.append((S.make_params(**c0_params(pulling=1, fixedEnd=0, pulledEnd=n_ramp - 1, deltax=0.02,
                                           pullFreq=100, seed=6)), s))
    jobs.append((S.make_params(**hi_kw(temperature=0.6, seed=7)), numpy_chain(n_tea, seed=83)))
    out = []
    for graph in (True, False):
        with S.Engine(0, 1 << 30) as e:
            e.set_chol_mode(mode)
            e.set_tuning(0, graph)
            ids = [e.add_job(p, sysm, 150) for p, sysm in jobs]
            assert e.block_steps() == 10
            e.run_blocks(15)
            e.synchronize()
            out.append([e.coords(j) for j in ids])
    for a, b in zip(*out):
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, b)
