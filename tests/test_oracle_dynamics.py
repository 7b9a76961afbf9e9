"""Pins the oracle's integrator and cadence: Euler closed form on a FENE
dimer, the reference's step count, the temperature accumulator, the
force-ramp cantilever and FP32-vs-FP64 zero-temperature trajectories."""
import numpy as np
import pytest

import oracle as O
from helpers import c0_params, numpy_chain
from sopgpu import System


def dimer(r=4.6):
    c = np.zeros((2, 4), np.float32)
    c[1, 0] = r
    e = np.zeros((0, 2), np.int32)
    z = np.zeros(0, np.float32)
    return System(c, np.array([[0, 1]], np.int32), np.array([3.8], np.float32), e, z, z)


def test_step_count_is_numsteps_plus_stride():
    s = O.Sim(O.make_params(**c0_params(temperature=0.0)), numpy_chain(30, seed=1))
    assert s.stride() == 10
    assert s.run(95) == 100  # blocks 0..9 of 10 steps
    assert s.step() == 100


def test_euler_closed_form_dimer():
    k, R0, h, zeta = 14.0, 2.0, 0.16, 50.0
    s = O.Sim(O.make_params(kspring_cov=k, R_limit=R0, timestep=h, zeta=zeta, temperature=0.0,
                            pairs_freq=10, possiblepairs_freq=10, nav=10, outputtiming=10, dcdfreq=10), dimer())
    s.run(190)  # 200 steps
    x0, x1 = 0.0, 4.6
    for _ in range(200):
        r = x1 - x0
        dr = r - 3.8
        f = k * dr / (1 - dr * dr / (R0 * R0))  # on bead 0 along +x
        x0 += h / zeta * f
        x1 -= h / zeta * f
    c = s.coords()
    assert c[0, 0] == pytest.approx(x0, abs=2e-5)
    assert c[1, 0] == pytest.approx(x1, abs=2e-5)


def test_temperature_accumulator_reset_and_value():
    s = O.Sim(O.make_params(**c0_params(nav=20, outputtiming=20, dcdfreq=20)), numpy_chain(50, seed=2))
    s.run(35)  # 40 steps: reset at step 0 and 20 → T holds steps 20..39
    T = s.temperature()
    assert np.all(T > 0)
    # T accumulates (zeta/6h)|dx|^2 per step; its mean over 20 steps and beads
    # is of the order of the temperature (0.6 kcal/mol)
    assert 0.2 < T.sum() / (50 * 20) < 3.0


def test_pulling_cantilever_and_fixed_bead():
    n = 40
    system = numpy_chain(n, seed=3)
    system.fixed = np.array([0], np.int32)
    system.pulled = np.array([n - 1], np.int32)
    kw = c0_params(temperature=0.0, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.5, pullFreq=10, k_trans=0.05)
    s = O.Sim(O.make_params(**kw), system)
    c0 = s.coords()
    s.run(25)  # 30 steps; updater at steps 0, 10, 20
    c = s.coords()
    assert np.array_equal(c[0], c0[0])  # fixed bead never moves
    ext = s.ext()
    # the last update (step 20) used the cantilever set at step 10: x0 + 0.5*1*v
    v = (c0[n - 1, :3] - c0[0, :3]).astype(np.float32)
    v = v / np.float32(np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
    assert ext[n - 1, 3] == 2.0
    assert np.linalg.norm(ext[n - 1, :3]) > 0


def test_zero_temperature_trajectory_fp32_vs_fp64():
    n = 200
    system = numpy_chain(n, seed=4)
    system.fixed = np.array([0], np.int32)
    system.pulled = np.array([n - 1], np.int32)
    kw = c0_params(temperature=0.0, pulling=1, fixedEnd=0, pulledEnd=n - 1, deltax=0.05, pullFreq=10, k_trans=0.05)
    s32 = O.Sim(O.make_params(**kw), system)
    s32.run(990)  # 1000 steps
    s64 = O.Sim(O.make_params(**kw), system)
    x64 = s64.run64(990)
    x32 = s32.coords()[:, :3].astype(np.float64)
    disp = x64 - system.coords[:, :3].astype(np.float64)
    err = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    assert err < 1e-5, err
    # the error is small against the motion itself, too
    assert np.linalg.norm(x32 - x64) < 1e-2 * np.linalg.norm(disp)
