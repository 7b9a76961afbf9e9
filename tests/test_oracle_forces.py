"""Pins the oracle's force and energy restatement with what the mathematics
fixes: two-bead closed forms, F = -grad U of the FP64 energies, Newton's
third law, and FP32-vs-FP64 agreement."""
import numpy as np
import pytest

import oracle as O
from helpers import numpy_chain
from sopgpu import System


def two_beads(r, r0=3.8, kind="bond", eh=1.2):
    c = np.zeros((2, 4), np.float32)
    c[1, 0] = r
    e = np.zeros((0, 2), np.int32)
    if kind == "bond":
        return System(c, np.array([[0, 1]], np.int32), np.array([r0], np.float32), e, np.zeros(0, np.float32),
                      np.zeros(0, np.float32))
    if kind == "native":
        return System(c, e, np.zeros(0, np.float32), np.array([[0, 1]], np.int32), np.array([r0], np.float32),
                      np.array([eh], np.float32))
    return System(c, e, np.zeros(0, np.float32), e, np.zeros(0, np.float32), np.zeros(0, np.float32),
                  pairs=np.array([[0, 1]], np.int32))


def sim(system, **kw):
    return O.Sim(O.make_params(**kw), system)


@pytest.mark.parametrize("r", [3.0, 3.8, 4.5, 5.5])
def test_fene_two_beads(r):
    k, R0, r0 = 14.0, 2.0, 3.8
    s = sim(two_beads(r, r0), kspring_cov=k, R_limit=R0)
    f = s.forces32()
    dr = r - r0
    # U = -(k/2) R0^2 log(1 - (r-r0)^2/R0^2); F_0 = -dU/dx_0 = k dr / (1 - dr^2/R0^2) along +x
    expect = k * dr / (1 - dr * dr / (R0 * R0))
    assert f[0, 0] == pytest.approx(expect, rel=1e-5, abs=1e-6)
    assert f[1, 0] == pytest.approx(-expect, rel=1e-5, abs=1e-6)
    if r == r0:
        assert f[0, 0] == 0.0


@pytest.mark.parametrize("r", [3.5, 5.0, 5.2, 7.0])
def test_native_two_beads(r):
    r0, eh = 5.0, 1.2
    s = sim(two_beads(r, r0, "native", eh))
    f = s.forces32()
    # U = eh((r0/r)^12 - 2 (r0/r)^6); F_0 = dU/dr (pointing to +x when attractive)
    dU = eh * (-12 * r0 ** 12 / r ** 13 + 12 * r0 ** 6 / r ** 7)
    assert f[0, 0] == pytest.approx(dU, rel=2e-5, abs=1e-6)
    assert f[1, 0] == pytest.approx(-dU, rel=2e-5, abs=1e-6)


@pytest.mark.parametrize("r", [3.0, 4.0, 10.0, 19.9, 20.1])
def test_repulsive_two_beads(r):
    a, el, rc = 3.8, 1.0, 20.0
    s = sim(two_beads(r, kind="pair"), a=a, el=el, pairs_cutoff=rc, pairlist_cutoff=rc, pairs_threshold=40.0)
    s.build_verlet()
    f = s.forces32()
    expect = -6 * el * a ** 6 / r ** 7 if r < rc else 0.0  # on bead 0, pushes to -x
    assert f[0, 0] == pytest.approx(expect, rel=1e-5, abs=1e-12)


def _energy64(s):
    return 0.5 * s.energies64().sum()


def test_force_is_minus_gradient():
    system = numpy_chain(40, seed=3)
    system.pairs = np.array([[i, j] for i in range(40) for j in range(i + 3, 40) if (i + j) % 3 == 0], np.int32)
    s = sim(system, kspring_cov=14.0)
    s.build_verlet()
    f64 = s.forces64()
    c0 = s.coords()
    h = 1e-3
    rng = np.random.default_rng(0)
    for bead in rng.choice(40, 8, replace=False):
        for d in range(3):
            cp = c0.copy(); cp[bead, d] += h
            cm = c0.copy(); cm[bead, d] -= h
            s.set_coords(cp); ep = _energy64(s)
            s.set_coords(cm); em = _energy64(s)
            s.set_coords(c0)
            fd = -(ep - em) / (2 * h)
            # coordinates are float32: h=1e-3 moves them by exact-enough amounts
            dx = float(cp[bead, d]) - float(cm[bead, d])
            fd = -(ep - em) / dx
            assert fd == pytest.approx(f64[bead, d], rel=2e-3, abs=2e-3)


def test_newton_third_law_and_fp32_vs_fp64():
    system = numpy_chain(200, seed=5)
    s = sim(system, kspring_cov=14.0, pairs_threshold=32.0)
    s.build_possible()
    s.build_verlet()
    f64 = s.forces64()
    f32 = s.forces32()[:, :3].astype(np.float64)
    assert np.abs(f64.sum(axis=0)).max() < 1e-9 * np.abs(f64).sum()
    err = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    assert err < 1e-5, err


def test_energy_sums_halve_pair_double_count():
    system = numpy_chain(60, seed=8)
    s = sim(system)
    s.build_verlet()
    e32 = s.energies32()
    sums = s.sum_energies()
    for p in range(3):
        assert sums[p, 0] == np.float32(np.float64(e32[p].astype(np.float64).sum()) * 0.5)
    # FENE energy at the native bond lengths is zero; native energy is -sum(eh)
    e64 = s.energies64()
    assert abs(0.5 * e64[0].sum()) < 1e-3
    assert 0.5 * e64[1].sum() == pytest.approx(-system.nat_eh.astype(np.float64).sum(), rel=1e-5)


def test_indentation_single_bead_closed_form():
    c = np.array([[0.0, 0.0, 70.0, 0.0]], np.float32)
    e = np.zeros((0, 2), np.int32)
    z = np.zeros(0, np.float32)
    system = System(c, e, z, e, z, z)
    kw = dict(indentation=1, indTip=(0, 0, 0), indChip=(0, 0, 0), indDirection=(0, 0, 1), indSurfR0=(0, 0, 100.0),
              indSurfN=(0, 0, -1), indTipR=50.0, indTipSigma=1.0, indTipEl=1.0, indTipA=0.0, indTipB=1.0,
              indSurfSigma=2.0, indSurfEl=0.5, indSurfA=1.0, indSurfB=-2.0)
    s = sim(system, **kw)
    f = s.forces32()
    d_tip = 70.0 - 50.0
    f_tip = 6.0 / d_tip ** 7  # U = (sigma/d)^6, pushes away from the tip (+z)
    h = 100.0 - 70.0  # distance to the plane along its normal (-z)
    A, B = 12 * 0.5 * 2.0 ** 12, 6 * 0.5 * 2.0 ** 6 * -2.0
    # the reference's plane term is (A'/h^12 + B'/h^6)/h^2 along the normal
    # (indentation_kernel.cu:111-119), i.e. -dU/dh divided by h; kept as is
    f_surf = (A / h ** 12 + B / h ** 6) / h ** 2
    assert f[0, 2] == pytest.approx(f_tip - f_surf, rel=1e-5)
    # the tip term alone is the exact gradient of (sigma/(rho-R))^6
    assert f_tip == pytest.approx(-(-6.0 / d_tip ** 7), rel=1e-12)
    assert abs(f[0, 0]) < 1e-12 and abs(f[0, 1]) < 1e-12
