"""Pins of the oracle's protocol updaters and hydrodynamic integrators
against their definitions, evaluated independently in FP64 numpy:

* force-ramp cantilever (src/Potentials/pulling.cu:152-165,198-214): at
  each updater pass the pulled bead gets k_trans (c_prev - x_pulled), with
  c_prev the cantilever set at the previous pass, c(s) = c0 + deltax
  floor(s / pullFreq) n;
* indentation tip (src/Potentials/indentation.cu:414-511): the tip moves by
  h nav / zeta_tip (Ks (chip - tip) - f_avg), with f_avg the wall's reaction
  averaged over nav steps, chip = chip0 + dir floor(s / nav) V, V reversed at
  the retraction step, optionally projected on dir;
* heating (src/Integrators/langevin.cu:68-82): T(s) = T0 + floor(s /
  tempFreq) dT at passes with s % tempFreq == 0, noise amplitude
  sqrt(2 T zeta / h);
* the TEA step (src/Integrators/bdhitea_kernel.cu:364-515) with its
  cross-term sum_j D_ij (F_j + beta' C_i rho g_j), listed and unlisted;
* the exact path's D F (bdhitea_kernel.cu:89-152) and its noise B rho g
  (bdhitea_kernel.cu:303-361) for the factor L itself and for the
  reference's reading of L's 3x3 blocks as symmetric matrices.

Each check is of one block of steps from a known state, so a dropped term,
a wrong sign, a wrong bead or an off-by-one cadence moves the result far
outside the FP32 rounding the tolerance allows."""
import numpy as np
import pytest

import oracle as O
from helpers import c0_params, numpy_chain
from sopgpu import System


def _free_beads(xyz):
    """Beads with no bonds and no contacts at the given positions."""
    c = np.zeros((len(xyz), 4), np.float32)
    c[:, :3] = np.asarray(xyz, np.float32)
    e = np.zeros((0, 2), np.int32)
    z = np.zeros(0, np.float32)
    return System(c, e, z, e, z, z)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / np.linalg.norm(b)


# ---------------------------------------------------------------------------
# force-ramp cantilever
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ntr", [1, 2])
def test_cantilever_force_uses_the_previous_cantilever_position(ntr):
    """Each trajectory pulls with its own cantilever (cantCoord[traj],
    pulling.cu:66-85,152-165): in a many-runs set the second trajectory's
    force still uses its own previous position, not one the first
    trajectory's pass has just advanced."""
    n = 40
    sysm = numpy_chain(n, seed=3)
    sysm.fixed = np.array([0], np.int32)
    sysm.pulled = np.array([n - 1], np.int32)
    k_trans, deltax, freq = 0.05, 0.5, 20
    kw = c0_params(temperature=0.0, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=deltax, pullFreq=freq,
                   k_trans=k_trans)
    s = O.Sim(O.make_params(**kw), sysm, ntr=ntr)
    stride = s.stride()
    assert stride == 10  # passes every 10 steps, the ramp every 20
    x0 = sysm.coords[:, :3].astype(np.float32)
    v = x0[n - 1] - x0[0]
    nhat = (v / np.float32(np.sqrt(np.float32(v @ v)))).astype(np.float64)
    cant0 = x0[n - 1].astype(np.float64)
    cant = cant0.copy()  # the position set by the previous pass
    rows = []
    for blk in range(8):
        step = blk * stride
        xc = s.coords()[:, :3].astype(np.float64)
        s.run(step)  # one block: the pass at `step`, then `stride` steps
        ext = s.ext()
        for t in range(ntr):
            xp = xc[t * n + n - 1]  # x_pulled of trajectory t at the pass
            f = ext[t * n + n - 1]
            assert f[3] == 2.0
            expect = k_trans * (cant - xp)
            np.testing.assert_allclose(f[:3], expect, rtol=2e-6, atol=1e-7, err_msg="trajectory %d" % t)
            if step % freq == 0:  # a pull .dat row: end-to-end, its projection, F . n, F
                e2e = xp - xc[t * n]
                rows.append([t, step, np.linalg.norm(e2e), e2e @ nhat, expect @ nhat, *expect])
        cant = cant0 + deltax * (step // freq) * nhat
    got = s.pull_rows()
    order = np.lexsort((got[:, 1], got[:, 0]))
    want = np.array(rows)
    np.testing.assert_allclose(got[order], want[np.lexsort((want[:, 1], want[:, 0]))], rtol=1e-5, atol=1e-6)
    # the ramp has moved the cantilever 3 deltax by step 70
    assert np.linalg.norm(cant - cant0) == pytest.approx(3 * deltax, rel=1e-6)


# ---------------------------------------------------------------------------
# indentation tip
# ---------------------------------------------------------------------------
def _tip_wall_force(bead, tip, R, Ap, Bp):
    """Reaction on the tip from one bead: (A'/(r-R)^12 + B'/(r-R)^6) /
    (r (r-R)) (bead - tip), r = |bead - tip| (indentation_kernel.cu:71-100)."""
    d = bead - tip
    r = np.linalg.norm(d)
    h = r - R
    return (Ap * h ** -12 + Bp * h ** -6) / (r * h) * d


@pytest.mark.parametrize("fix_trans,retract", [(0, -1), (1, 40)])
def test_tip_update_closed_form(fix_trans, retract):
    # two fixed beads (the reaction on the tip is then constant between
    # passes), the tip 1.3 A off its wall from the first, the chip above
    # moving down
    beads = [(0.0, 0.0, 0.0), (150.0, 0.0, 0.0)]
    sysm = _free_beads(beads)
    sysm.fixed = np.array([0, 1], np.int32)
    R, Ks, zeta_tip, V, nav, h = 5.0, 0.5, 40.0, 0.3, 10, 0.16
    sigma, el, A, B = 1.2, 0.8, 0.5, 1.0
    tip0 = np.array([1.0, 1.5, 6.0])
    chip0 = np.array([0.0, 0.0, 12.0])
    direction = np.array([0.3, 0.0, -1.0])
    kw = c0_params(temperature=0.0, pulling=1, fconst=1.0, fixedEnd=0, pulledEnd=1, indentation=1,
                   indTip=tuple(tip0), indChip=tuple(chip0), indDirection=tuple(direction),
                   indSurfR0=(0.0, 0.0, -50.0), indSurfN=(0.0, 0.0, 1.0), indTipR=R, indKs=Ks, indDeltaX=V,
                   indTipZeta=zeta_tip, indTipSigma=sigma, indTipEl=el, indTipA=A, indTipB=B,
                   indFixTrans=fix_trans, indRetractionStep=retract, nav=nav, outputtiming=nav, dcdfreq=nav,
                   pairs_freq=nav, possiblepairs_freq=nav, timestep=h, indOutputFreq=2 * nav)
    s = O.Sim(O.make_params(**kw), sysm)
    assert s.stride() == nav
    Ap, Bp = 12 * el * sigma ** 12 * A, 6 * el * sigma ** 6 * B
    dirn = direction / np.linalg.norm(direction)
    tip, vel = tip0.copy(), V
    mult = h * nav / zeta_tip
    f_prev = np.zeros(3)  # reaction accumulated before the first pass: none
    acc, rows = [], []
    for blk in range(8):
        step = blk * nav
        if step == retract:
            vel = -vel
        # the pass at `step`: average of the previous nav steps' reaction
        chip = chip0 + dirn * (step // nav) * vel
        df = -f_prev + Ks * (chip - tip)
        kdr = Ks * np.linalg.norm(chip - tip)
        tip = tip + (mult * (df @ dirn)) * dirn if fix_trans else tip + mult * df
        # indentation .dat: averages over the passes since the last row, one
        # row every indentationOutputFreq = 2 nav steps
        acc.append((f_prev, np.linalg.norm(f_prev), kdr))
        if step % (2 * nav) == 0:
            fa = sum(a[0] for a in acc) / 2
            rows.append([step, (step // nav) * vel, sum(a[1] for a in acc) / 2, np.linalg.norm(fa), kdr,
                         sum(a[2] for a in acc) / 2, *fa, *tip, *chip])
            acc = []
        s.run(step)
        ind = s.indentation()
        np.testing.assert_allclose(ind[3:6], chip, rtol=1e-6, atol=1e-5)
        np.testing.assert_allclose(ind[0:3], tip, rtol=1e-5, atol=1e-5)
        # the block's nav steps each add the reaction at this tip position
        f_prev = sum(_tip_wall_force(np.array(b), tip, R, Ap, Bp) for b in beads)
    assert np.linalg.norm(tip - tip0) > 0.5  # the tip did move
    got = s.ind_rows()
    assert len(got) == len(rows) == 4
    np.testing.assert_allclose(got, np.array(rows), rtol=2e-5, atol=2e-5)


# ---------------------------------------------------------------------------
# heating
# ---------------------------------------------------------------------------
def test_heating_schedule_sets_the_noise_amplitude():
    # more than pairs_threshold (32 A) apart, close to the origin so the
    # coordinates' rounding stays far below the per-block displacement
    xyz = [(0.0, 0.0, 0.0), (40.0, 0.0, 0.0), (0.0, 40.0, 0.0)]
    sysm = _free_beads(xyz)
    T0, dT, freq, h, zeta = 0.2, 0.15, 30, 0.16, 50.0
    kw = c0_params(heating=1, initialT=T0, deltaT=dT, tempFreq=freq, nav=10, outputtiming=10, dcdfreq=10,
                   pairs_freq=10, possiblepairs_freq=10, timestep=h, zeta=zeta, seed=99)
    s = O.Sim(O.make_params(**kw), sysm)
    assert s.stride() == 10
    state = s.rng()
    for blk in range(12):
        step = blk * 10
        x_before = s.coords()[:, :3].astype(np.float64)
        s.run(step)
        g, state = O.gaussians(state, 10)  # the block's draws, bead by bead
        dx = s.coords()[:, :3].astype(np.float64) - x_before
        # free beads: x advances by (h/zeta) rho g every step, rho fixed in
        # the block; rho by least squares over the nine components
        G = (h / zeta) * g.astype(np.float64).sum(axis=1).ravel()
        rho = float(dx.ravel() @ G / (G @ G))
        T = T0 + (step // freq) * dT
        assert rho == pytest.approx(np.sqrt(2 * T * zeta / h), rel=1e-5), (step, rho)
        np.testing.assert_allclose(dx.ravel(), rho * G, rtol=0, atol=2e-5)


# ---------------------------------------------------------------------------
# hydrodynamics: one step from a known state
# ---------------------------------------------------------------------------
def _one_step_kw(**kw):
    d = c0_params(nav=1, outputtiming=1, dcdfreq=1, pairs_freq=1, possiblepairs_freq=1, hi_on=1, hi_a=1.9,
                  hi_epsilon_freq=1, hi_capricious=0, seed=2718)
    d.update(kw)
    return d


def _step_inputs(sysm, kw):
    """Forces (FP64 oracle, pinned by test_oracle_forces), noise draws and
    the first step's displacement of the oracle."""
    p = O.make_params(**kw)
    f = O.Sim(p, sysm)
    f.build_possible()
    f.build_verlet()
    F = f.forces64()
    lists = [f.list(w) for w in (0, 1, 2)]
    s = O.Sim(O.make_params(**kw), sysm)
    assert s.stride() == 1
    g, _ = O.gaussians(s.rng(), 1)
    x0 = s.coords()[:, :3].astype(np.float64)
    assert s.run(0) == 1
    dx = s.coords()[:, :3].astype(np.float64) - x0
    rho = np.sqrt(2 * kw["temperature"] * kw["zeta"] / kw["timestep"])
    # each coordinate is stored in FP32, so the displacement read back
    # carries up to an ulp of the coordinate's magnitude
    ulp = float(np.spacing(np.float32(np.abs(x0).max() + np.abs(dx).max())))
    return F, rho * g[:, 0, :].astype(np.float64), dx, lists, ulp


def _blocks(D, n):
    return D.reshape(n, 3, n, 3).transpose(0, 2, 1, 3)  # [i, j] -> 3x3 block D_ij


@pytest.mark.parametrize("unlisted", [1, 0])
def test_tea_step_with_cross_term(unlisted):
    n = 60
    sysm = numpy_chain(n, seed=21)
    kw = _one_step_kw(temperature=0.6, hi_unlisted=unlisted)
    F, noise, dx, lists, ulp = _step_inputs(sysm, kw)
    D = O.rpy_matrix(sysm.coords, kw["hi_a"])
    Db = _blocks(D, n)
    partners = []
    for i in range(n):
        if unlisted:
            partners.append([j for j in range(n) if j != i])
        else:
            row = []
            for cnt, ent in lists:
                row += list(ent[: cnt[i], i])
            partners.append(row)
    # epsilon: mean of the off-diagonal-block entries over the partners the
    # update visits, beta' from its closed form (bdhitea.cu:145-170)
    tot = sum(Db[i, j].sum() for i in range(n) for j in partners[i])
    eps = tot / (3 * n * (3 * n - 3))
    a = (3 * n - 1) * eps ** 2 - (3 * n - 2) * eps
    beta = (1 - np.sqrt(1 - a)) / a
    expect = np.zeros((n, 3))
    for i in range(n):
        sq = sum((Db[i, j] ** 2).sum(axis=1) for j in partners[i])  # row sums of squares
        C = 1.0 / np.sqrt(1.0 + beta ** 2 * sq)
        f = F[i] + C * noise[i]
        for j in partners[i]:
            f = f + Db[i, j] @ (F[j] + beta * C * noise[j])
        expect[i] = kw["timestep"] / kw["zeta"] * f
    assert _rel(dx, expect) < 2e-5
    np.testing.assert_allclose(dx, expect, rtol=2e-5, atol=2 * ulp)
    # the cross-term is no rounding-level correction
    assert _rel(dx, kw["timestep"] / kw["zeta"] * (F + noise)) > 0.05


def test_exact_path_zero_temperature_is_D_times_F():
    n = 50
    sysm = numpy_chain(n, seed=22)
    # low friction so the step (~1e-2 A) is large against the coordinates' ulp
    kw = _one_step_kw(temperature=0.0, hi_exact=1, zeta=2.0)
    F, _, dx, _, ulp = _step_inputs(sysm, kw)
    D = O.rpy_matrix(sysm.coords, kw["hi_a"])
    expect = kw["timestep"] / kw["zeta"] * (D @ F.ravel()).reshape(n, 3)
    assert _rel(dx, expect) < 1e-4
    np.testing.assert_allclose(dx, expect, rtol=2e-5, atol=2 * ulp)
    assert _rel(dx, kw["timestep"] / kw["zeta"] * F) > 0.05


@pytest.mark.parametrize("faithful", [0, 1])
def test_exact_path_noise_is_the_factor_times_g(faithful):
    n = 50
    sysm = numpy_chain(n, seed=23)
    kw = _one_step_kw(temperature=0.6, hi_exact=1, hi_faithful_blocks=faithful)
    F, noise, dx, _, ulp = _step_inputs(sysm, kw)
    D = O.rpy_matrix(sysm.coords, kw["hi_a"])
    L = np.linalg.cholesky(D)
    if faithful:
        # bdhitea_kernel.cu:322-337 reads, for block (i, j <= i), the six
        # entries of L's block on and below its diagonal and applies them as
        # a symmetric matrix
        B = np.zeros_like(L)
        for i in range(n):
            for j in range(i + 1):
                blk = L[3 * i:3 * i + 3, 3 * j:3 * j + 3]
                B[3 * i:3 * i + 3, 3 * j:3 * j + 3] = np.tril(blk) + np.tril(blk, -1).T
    else:
        B = L
    expect = kw["timestep"] / kw["zeta"] * ((D @ F.ravel()) + B @ noise.ravel()).reshape(n, 3)
    assert _rel(dx, expect) < 2e-5
    np.testing.assert_allclose(dx, expect, rtol=2e-5, atol=2 * ulp)
    other = np.linalg.cholesky(D) if faithful else B  # the other reading differs visibly
    if faithful:
        alt = kw["timestep"] / kw["zeta"] * ((D @ F.ravel()) + other @ noise.ravel()).reshape(n, 3)
        assert _rel(dx, alt) > 1e-3
