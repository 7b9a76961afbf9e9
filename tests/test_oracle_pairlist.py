"""Pins the oracle's possible-pairs and Verlet builds against the set
definition written independently in numpy (float32 arithmetic, no FMA)."""
import numpy as np

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
