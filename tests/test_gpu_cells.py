"""Cell-binned possible pairs on the GPU.  The reference builds the list by
scanning every same-trajectory bead (src/Updaters/possiblepairlist_kernel.cu:18-71);
jobs of 8192 beads and more here take their candidates from a cell grid.
The list -- entries, their ascending order and the counts -- must be the
one the all-pairs scan gives: checked against the oracle on a 20,000-bead
assembly, and against the engine's own all-pairs scan on the 100,000-bead
capsid shell (200 chains x 500 beads, R = 140 A) at step 0 and after
10^4 steps of continuous indentation."""
import numpy as np
import pytest

import oracle as O
import sopgpu as S
from helpers import c0_params

pytestmark = pytest.mark.gpu


def lists_equal(a_cnt, a_ent, b_cnt, b_ent):
    if not np.array_equal(a_cnt, b_cnt):
        return False
    m = int(a_cnt.max()) if len(a_cnt) else 0
    ka = np.arange(a_ent.shape[0])[:, None] < a_cnt[None, :]
    kb = np.arange(b_ent.shape[0])[:, None] < b_cnt[None, :]
    return np.array_equal(np.where(ka, a_ent, -1)[:m], np.where(kb, b_ent, -1)[:m])


def test_cells_match_oracle_20k():
    sysm = S.synth_shell(40, 500, seed=5, radius=80.0)
    assert sysm.n == 20000
    kw = c0_params(pairs_threshold=32.0, max_possiblePairs=1024)
    eng = S.Engine(0, 8 << 30)
    try:
        print("case: 20k shell, cells vs oracle at step 0", flush=True)
        jid = eng.add_job(S.make_params(**kw), sysm, 0)
        eng.rebuild(jid, 3)
        assert eng.status(jid)[2] == 0
        ref = O.Sim(O.make_params(**kw), sysm)
        ref.build_possible()
        ref.build_verlet()
        assert lists_equal(*eng.list(jid, 3), *ref.list(3))
        assert lists_equal(*eng.list(jid, 2), *ref.list(2))
    finally:
        eng.close()


def test_cells_trajectory_equals_all_pairs_scan():
    sysm = S.synth_shell(40, 500, seed=6, radius=80.0)
    kw = c0_params(pairs_threshold=32.0, max_possiblePairs=1024, temperature=0.6)
    out = []
    for cells in (True, False):
        print("case: 20k shell, 300 steps, cells=%s" % cells, flush=True)
        eng = S.Engine(0, 8 << 30)
        try:
            eng.set_cells(enabled=cells)
            jid = eng.add_job(S.make_params(**kw), sysm, 290)
            eng.run_job(jid)
            assert eng.status(jid)[0] == 300
            out.append((eng.coords(jid), eng.list(jid, 3), eng.list(jid, 2)))
        finally:
            eng.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert lists_equal(*out[0][1], *out[1][1])
    assert lists_equal(*out[0][2], *out[1][2])


@pytest.fixture(scope="module")
def capsid():
    return S.synth_shell(200, 500, seed=2026, radius=140.0)


def capsid_params(sysm):
    c = sysm.coords[:, :3]
    top, bot = float(c[:, 2].max()), float(c[:, 2].min())
    tip_r = 50.0
    # tip 1 um/s in 0.16 tau_H (40 ps) steps: 4e-4 A per step, 0.04 A per nav=100 update
    return c0_params(pairs_threshold=32.0, max_possiblePairs=1024, temperature=0.6, nav=100,
                     outputtiming=100, dcdfreq=100, indentation=1, indTip=(0.0, 0.0, top + tip_r + 4.0),
                     indChip=(0.0, 0.0, top + tip_r + 4.0), indDirection=(0.0, 0.0, -1.0),
                     indSurfR0=(0.0, 0.0, bot - 4.0), indSurfN=(0.0, 0.0, 1.0), indTipR=tip_r, indKs=0.07,
                     indDeltaX=0.04, indTipZeta=5000.0, indTipA=0.0, indTipB=1.0)


def test_capsid_100k_lists_at_steps_0_and_10000(capsid):
    assert capsid.n == 100000
    kw = capsid_params(capsid)
    eng = S.Engine(0, 16 << 30)
    try:
        jid = eng.add_job(S.make_params(**kw), capsid, 9990)
        for label in ("step 0", "step 10000"):
            print("case: capsid 100k, lists at %s" % label, flush=True)
            eng.set_cells(enabled=True)
            eng.rebuild(jid, 3)
            a3, a2 = eng.list(jid, 3), eng.list(jid, 2)
            eng.set_cells(enabled=False)
            eng.rebuild(jid, 3)
            b3, b2 = eng.list(jid, 3), eng.list(jid, 2)
            assert eng.status(jid)[2] == 0
            assert lists_equal(*a3, *b3)
            assert lists_equal(*a2, *b2)
            if label == "step 0":
                eng.set_cells(enabled=True)
                print("case: capsid 100k, 10^4 steps with cells", flush=True)
                eng.run_job(jid)
                assert eng.status(jid)[0] == 10000
        c = eng.coords(jid)
        assert np.isfinite(c).all()
    finally:
        eng.close()
