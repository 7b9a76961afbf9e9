"""Pins the oracle's possible-pairs and Verlet builds against the set
definition written independently in numpy (float32 arithmetic, no FMA)."""
import numpy as np
import pytest

import oracle as O
from helpers import numpy_chain


def np_dist32(c, i, j):
    d = c[j, :3] - c[i, :3]  # float32, (j - i) as the kernels form it
    s = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    s = s + d[..., 2] * d[..., 2]
    return np.sqrt(s)


def brute_possible(system, thr, ntr=1):
    c = np.tile(system.coords, (ntr, 1)).astype(np.float32)
    N = system.n
    n = N * ntr
    excl = [set() for _ in range(N)]
    for a, b in system.bonds:
        excl[a].add(b); excl[b].add(a)
    for a, b in system.natives:
        excl[a].add(b); excl[b].add(a)
    out = []
    for i in range(n):
        t = i // N
        js = np.arange(t * N, (t + 1) * N)
        d = np_dist32(c, i, js)
        keep = [int(j) for j, dj in zip(js, d) if j != i and dj < np.float32(thr) and (j - t * N) not in excl[i - t * N]]
        out.append(keep)
    return out


def lists_of(cnt, ent):
    return [list(ent[: cnt[i], i]) for i in range(len(cnt))]


def test_possible_pairs_match_brute_force():
    system = numpy_chain(150, seed=21)
    s = O.Sim(O.make_params(pairs_threshold=16.0, pairlist_cutoff=10.0), system)
    s.build_possible()
    assert lists_of(*s.list(3)) == brute_possible(system, 16.0)


def test_possible_pairs_many_runs():
    system = numpy_chain(60, seed=22)
    s = O.Sim(O.make_params(pairs_threshold=12.0), system, ntr=3)
    s.build_possible()
    assert lists_of(*s.list(3)) == brute_possible(system, 12.0, ntr=3)


def test_verlet_filters_in_order():
    system = numpy_chain(150, seed=23)
    s = O.Sim(O.make_params(pairs_threshold=16.0, pairlist_cutoff=9.0), system)
    s.build_possible()
    s.build_verlet()
    c = system.coords.astype(np.float32)
    pp = lists_of(*s.list(3))
    expect = [[j for j in pp[i] if np_dist32(c, i, np.array([j]))[0] < np.float32(9.0)] for i in range(150)]
    assert lists_of(*s.list(2)) == expect


def test_overflow_is_flagged():
    system = numpy_chain(80, seed=24)
    s = O.Sim(O.make_params(pairs_threshold=30.0, max_possiblePairs=5), system)
    s.build_possible()
    assert s.overflow() == 1


def _sqrt_below_bound(t):
    """Largest float32 x with sqrtf(x) < t, found as the engine finds it."""
    t = np.float32(t)
    x = np.float32(t * t)
    while x > 0 and not np.sqrt(x) < t:
        x = np.nextafter(x, np.float32(0))
    while True:
        y = np.nextafter(x, np.float32(np.inf))
        if not np.sqrt(y) < t:
            return x
        x = y


@pytest.mark.parametrize("t", [20.0, 32.0, 12.0, 7.3, 200.0, 0.5, 3.8, 1e-3])
def test_list_predicate_as_squared_bound(t):
    # The kernels test sqrtf(d2) < cut as d2 <= bound: sqrtf is correctly
    # rounded and monotone, so the two agree on every float d2.  Checked on
    # every float within 4000 ulps of the bound and on a random sample.
    b = _sqrt_below_bound(t)
    t32 = np.float32(t)
    bi = np.array([b], np.float32).view(np.int32)[0]
    window = np.arange(bi - 4000, bi + 4000, dtype=np.int32).view(np.float32)
    window = window[window >= 0]
    np.testing.assert_array_equal(np.sqrt(window) < t32, window <= b)
    rng = np.random.default_rng(1)
    d2 = (rng.random(200000, dtype=np.float32) * np.float32(4.0 * t * t)).astype(np.float32)
    np.testing.assert_array_equal(np.sqrt(d2) < t32, d2 <= b)
