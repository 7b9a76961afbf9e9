"""Pins the oracle's hydrodynamics: the RPY tensor against its closed form
and positive definiteness, TEA's epsilon and beta' against their
definitions on the FP64 tensor, the Cholesky restatement against LAPACK,
and the TEA limit of vanishing coupling."""
import numpy as np
import pytest

import oracle as O
from helpers import c0_params, numpy_chain


def hi_params(**kw):
    d = c0_params(hi_on=1, hi_a=1.9, hi_epsilon_freq=10, hi_capricious=0)
    d.update(kw)
    return O.make_params(**d)


def test_rpy_tensor_closed_form_and_spd():
    sysm = numpy_chain(60, seed=1)
    D = O.rpy_matrix(sysm.coords, 3.8)
    assert np.allclose(D, D.T)
    assert np.linalg.eigvalsh(D).min() > 0
    # one far and one overlapping pair by hand (Rotne-Prager-Yamakawa)
    c = sysm.coords[:, :3].astype(np.float64)
    for i, j in ((0, 40), (10, 11)):
        r = c[j] - c[i]
        d = np.linalg.norm(r)
        u = r / d
        ra = d / 3.8
        if ra > 2:
            blk = 0.75 / ra * ((1 + 2 / (3 * ra * ra)) * np.eye(3) + (1 - 2 / (ra * ra)) * np.outer(u, u))
        else:
            blk = (1 - 9 * ra / 32) * np.eye(3) + 3 * ra / 32 * np.outer(u, u)
        np.testing.assert_allclose(D[3 * i:3 * i + 3, 3 * j:3 * j + 3], blk, rtol=1e-12, atol=1e-14)


def test_tea_epsilon_and_beta_definition():
    n = 80
    sysm = numpy_chain(n, seed=2)
    s = O.Sim(hi_params(temperature=0.0), sysm)
    s.run(5)  # one TEA update at step 0, then 10 steps
    eps, beta = s.hi_tea()
    D = O.rpy_matrix(sysm.coords, 1.9)
    mask = np.ones_like(D, bool)
    for i in range(n):
        mask[3 * i:3 * i + 3, 3 * i:3 * i + 3] = False
    e_ref = D[mask].mean()
    assert eps[0] == pytest.approx(e_ref, rel=1e-5)
    N3 = 3 * n
    a = (N3 - 1) * e_ref ** 2 - (N3 - 2) * e_ref
    assert beta[0] == pytest.approx((1 - np.sqrt(1 - a)) / a, rel=1e-5)


def test_tea_vanishing_coupling_is_langevin():
    n = 50
    sysm = numpy_chain(n, seed=3)
    kw = dict(temperature=0.6, seed=77)
    plain = O.Sim(O.make_params(**c0_params(**kw)), sysm)
    plain.run(95)
    tea = O.Sim(hi_params(hi_a=1e-6, **kw), sysm)
    tea.run(95)
    # D_ij/D_ii ~ 0.75 a/r -> 0: same noise draws, same forces
    np.testing.assert_allclose(tea.coords(), plain.coords(), rtol=0, atol=2e-5)


def test_cholesky_restatement_vs_lapack():
    n = 40
    sysm = numpy_chain(n, seed=4)
    s = O.Sim(hi_params(hi_exact=1, hi_a=1.9, temperature=0.0), sysm)
    s.run(5)  # the last refactorisation is at step 9 from the step-9 coordinates
    L = s.hi_factor().astype(np.float64)
    c = s.coords()  # after step 10; rebuild D at the step-9 coordinates instead:
    s2 = O.Sim(hi_params(hi_exact=1, hi_a=1.9, temperature=0.0, hi_chol_freq=1000000), sysm)
    s2.run(5)
    L0 = s2.hi_factor().astype(np.float64)  # factor of the initial structure
    D0 = O.rpy_matrix(sysm.coords, 1.9)
    Lref = np.linalg.cholesky(D0)
    assert np.abs(np.triu(L0, 1)).max() == 0.0
    assert np.linalg.norm(L0 @ L0.T - D0) / np.linalg.norm(D0) < 2e-6
    assert np.linalg.norm(L0 - Lref) / np.linalg.norm(Lref) < 2e-5
    assert np.isfinite(L).all() and np.isfinite(c).all()


def test_exact_noise_covariance():
    """Noise sum_j B_ij rho g_j: with the factor applied as L it has
    covariance rho^2 D; the reference's block reading gives a different one."""
    n = 6
    sysm = numpy_chain(n, seed=5)
    D = O.rpy_matrix(sysm.coords, 1.9)
    Lref = np.linalg.cholesky(D)
    # symmetric-block reading (bdhitea_kernel.cu:322-337) of the exact factor
    B = np.zeros_like(Lref)
    for d in range(n):
        for i in range(d + 1):
            blk = Lref[3 * d:3 * d + 3, 3 * i:3 * i + 3]
            low = np.tril(blk)
            B[3 * d:3 * d + 3, 3 * i:3 * i + 3] = low + np.tril(blk, -1).T
    assert np.linalg.norm(B @ B.T - D) / np.linalg.norm(D) > 1e-3
    assert np.linalg.norm(Lref @ Lref.T - D) / np.linalg.norm(D) < 1e-12


@pytest.mark.parametrize("faithful", [0, 1])
def test_exact_hi_runs_and_stays_finite(faithful):
    n = 30
    sysm = numpy_chain(n, seed=6)
    s = O.Sim(hi_params(hi_exact=1, hi_faithful_blocks=faithful), sysm)
    s.run(45)
    assert np.isfinite(s.coords()).all()
