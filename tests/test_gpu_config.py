"""The sop-gpu program end to end on the GPU: a parameter file in the
reference's format, run through the C ABI (sop_run_config), against the
oracle's runGPU (src/gsop.cu:111-163) on the same topology, coordinates
and seeds.  Final PDB coordinates and the energy .dat rows must match the
oracle's text exactly; the other outputs are checked for the cadence and
layout the reference writes (src/Updaters/output_manager.cu:339-360,
src/Potentials/pulling.cu:166-187, src/Updaters/dcd_manager.cpp:70-107,
src/IO/pdbio.cpp:194-219)."""
import os

import numpy as np
import pytest

import oracle as O
import sopgpu as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOP = os.path.join(GOLDEN, "tiny.top")
PDB = os.path.join(GOLDEN, "tiny.pdb")

pytestmark = pytest.mark.gpu


def write_cfg(path, kv):
    with open(path, "w") as f:
        for k, v in kv.items():
            f.write("%s\t\t%s\n" % (k, v))


def atom_lines(path):
    with open(path) as f:
        return [line for line in f if line.startswith("ATOM")]


def xyz_text(path):
    return [line[30:54] for line in atom_lines(path)]


def oracle_xyz_text(coords):
    return ["%8.3f%8.3f%8.3f" % (float(c[0]), float(c[1]), float(c[2])) for c in coords]


def dat_rows(path):
    with open(path) as f:
        return [line.rstrip("\n").split("\t") for line in f]


def oracle_dat_text(row):
    e = [np.float32(row[2]), np.float32(row[3]), np.float32(row[4])]
    total = float(e[0]) + float(e[1]) + float(e[2])
    return ["%12d" % int(row[0]), "%8.5f" % float(np.float32(row[1]))] + ["%8.5f" % float(x) for x in e] + \
        ["%8.5f" % total, "%8d" % int(row[5])]


def pull_text(r):
    """A pull .dat row of the oracle, printed with the reference's format."""
    return "%12d\t%5.3f\t%5.3f\t%5.3f\t%5.3f\t%5.3f\t%5.3f" % (int(r[1]), *[float(np.float32(x)) for x in r[2:8]])


def ind_text(r):
    """An indentation .dat row of the oracle (indentation.cu:526-533)."""
    return "%d\t" % int(r[0]) + "\t".join("%8.5f" % float(np.float32(x)) for x in r[1:15])


def test_many_runs_pulling_matches_oracle(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    os.makedirs("dat")
    n, steps, seed, firstrun = 40, 1200, 77, 3
    cfg = {
        "name": "tiny", "topology": TOP, "coordinates": PDB, "runnum": 2, "firstrun": firstrun,
        "numsteps": steps, "timestep": 0.16, "seed": seed, "temperature": 0.6, "nav": 100,
        "outputtiming": 200, "dcdfreq": 400, "restartfreq": 600, "pairs_freq": 10, "possiblepairs_freq": 100,
        "pairs_threshold": 32.0, "kspring_cov": 14.0, "pulling": "on", "fixed": "0", "pulled": "23",
        "fixedEnd": 0, "pulledEnd": 23, "deltax": 0.05, "k_trans": 0.05, "pullFreq": 100,
        "outputname": "dat/<name>_<run>.dat", "pullOutput": "pull_<run>.dat", "DCDfile": "<name>_<run>.dcd",
        "finalcoord": "<name>_<run>_final.pdb", "reffilename": "<name>.ref.pdb", "restartname": "restart_<run>.pdb",
    }
    write_cfg("run.sop", cfg)
    print("case: many runs + pulling, N=40, 2 runs, %d steps" % steps, flush=True)
    S.run_config("run.sop")

    sys_ = S.read_topology(TOP, PDB)
    sys_.fixed = np.array([0], np.int32)
    sys_.pulled = np.array([23], np.int32)
    p = O.make_params(kspring_cov=14.0, pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100,
                      timestep=0.16, temperature=0.6, seed=seed + firstrun, nav=100, outputtiming=200, dcdfreq=400,
                      pulling=1, deltax=0.05, k_trans=0.05, pullFreq=100, fixedEnd=0, pulledEnd=23)
    ref = O.Sim(p, sys_, ntr=2)
    ref.run(steps)
    rc = ref.coords()
    for t in range(2):
        assert xyz_text("tiny_%d_final.pdb" % (t + firstrun)) == oracle_xyz_text(rc[t * n:(t + 1) * n])
    # energy .dat of the first run: every column as the reference prints it
    rows = dat_rows("dat/tiny_%d.dat" % firstrun)
    want = ref.dat()
    assert len(rows) == len(want) == steps // 200 + 1
    for got, w in zip(rows, want):
        assert got[:7] == oracle_dat_text(w)
        assert got[7] == "%8.5f" % 0.0  # Rg column, computeRg off
    assert len(dat_rows("dat/tiny_%d.dat" % (firstrun + 1))) == len(want)
    # pull .dat (pulling.cu:166-187): every row as the reference prints it
    prows = ref.pull_rows()
    for t in range(2):
        with open("pull_%d.dat" % (t + firstrun)) as f:
            got = [line.rstrip("\n") for line in f]
        want = [pull_text(r) for r in prows if int(r[0]) == t]
        assert len(want) == steps // 100 + 1
        assert got == want
    # DCD: one frame per dcdfreq steps of 40 atoms
    size = os.path.getsize("tiny_%d.dcd" % firstrun)
    assert size == 276 + (steps // 400 + 1) * 3 * (4 * n + 8)
    # reference PDB: B-factors carry the fixed (1) / pulled (2) flags
    ref_atoms = atom_lines("tiny.ref.pdb")
    betas = [float(line[60:66]) for line in ref_atoms]
    assert betas[0] == 1.0 and betas[23] == 2.0 and sum(betas) == 3.0
    assert os.path.exists("restart_%d.pdb" % firstrun)


def test_single_run_tea_force_clamp_matches_oracle(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    n, steps, seed, run = 40, 400, 5032014, 5
    cfg = {
        "name": "tiny", "topology": TOP, "coordinates": PDB, "run": run,
        "numsteps": steps, "timestep": 0.16, "seed": seed, "temperature": 0.6, "nav": 100,
        "outputtiming": 100, "dcdfreq": 200, "pairs_freq": 10, "possiblepairs_freq": 100, "pairs_threshold": 32.0,
        "pulling": "on", "fixed": "0", "pulled": "39", "pullDirection": "vector", "pullVector": "1.0 0.0 0.0",
        "fconst": 1.43, "fixedEnd": 0, "pulledEnd": 39,
        "hi_on": "yes", "hi_a": 1.9, "hi_unlisted": "yes", "hi_capricious": "no", "hi_epsilon_freq": 10,
        "outputname": "<name>_<run>.dat", "finalcoord": "<name>_<run>_final.pdb",
    }
    write_cfg("run.sop", cfg)
    print("case: TEA + force clamp, N=40, run %d, %d steps" % (run, steps), flush=True)
    S.run_config("run.sop")
    sys_ = S.read_topology(TOP, PDB)
    sys_.fixed = np.array([0], np.int32)
    sys_.pulled = np.array([39], np.int32)
    p = O.make_params(pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=0.16, temperature=0.6,
                      seed=seed + run, nav=100, outputtiming=100, dcdfreq=200, pulling=1, fconst=1.43,
                      pullDirectionVector=1, pullVector=(1.0, 0.0, 0.0), fixedEnd=0, pulledEnd=39,
                      hi_on=1, hi_a=1.9, hi_unlisted=1, hi_capricious=0, hi_epsilon_freq=10)
    ref = O.Sim(p, sys_, ntr=1)
    ref.run(steps)
    # `run 5` masks <run> in file names with the parameter's own text
    assert xyz_text("tiny_%d_final.pdb" % run) == oracle_xyz_text(ref.coords())
    rows = dat_rows("tiny_%d.dat" % run)
    want = ref.dat()
    assert len(rows) == len(want)
    for got, w in zip(rows, want):
        assert got[:7] == oracle_dat_text(w)
        assert len(got) == 9  # TEA epsilon column
    assert 0.0 < float(rows[-1][8]) < 1.0
    # no pull .dat for a force clamp (deltax = 0): the literal default name is never created
    assert not any(f.startswith("pull.") for f in os.listdir("."))


def test_indentation_outputs_match_oracle(tmp_path, monkeypatch):
    """Indentation of the 40-bead chain: final coordinates, energy .dat and
    indentation .dat rows identical as text to the oracle's; the tip,
    cantilever and surface representation beads of the reference PDB and
    the VMD connect script as indentation.cu:254-400 builds them."""
    monkeypatch.chdir(tmp_path)
    sys_ = S.read_topology(TOP, PDB)
    c = sys_.coords[:, :3].astype(np.float64)
    top, bot = float(c[:, 2].max()), float(c[:, 2].min())
    # positions as the parameter file states them (three decimals), which is
    # what both the program and the oracle then read
    tip = tuple(round(v, 3) for v in (float(c[:, 0].mean()), float(c[:, 1].mean()), top + 12.0))
    surf_z = round(bot - 6.0, 3)
    steps, seed, run = 1200, 4711, 2
    size, grid_step, cant_len = 3, 1.4, 500.0
    cfg = {
        "name": "tiny", "topology": TOP, "coordinates": PDB, "run": run, "numsteps": steps, "timestep": 0.16,
        "seed": seed, "temperature": 0.6, "nav": 100, "outputtiming": 100, "dcdfreq": 400, "pairs_freq": 10,
        "possiblepairs_freq": 100, "pairs_threshold": 32.0, "indentation": "on",
        "indentationChip": "%.3f %.3f %.3f" % tip, "indentationTip": "%.3f %.3f %.3f" % tip,
        "indentationDirection": "0.0 0.0 -1.0", "indentationSurfaceR0": "0.0 0.0 %.3f" % surf_z,
        "indentationSurfaceN": "0.0 0.0 1.0", "indentationTipR": 8.0, "indentationKs": 0.07,
        "indentationDeltaX": 0.8, "indentationTipZeta": 500.0, "indentationOutputFreq": 200,
        "indentationOutput": "ind_<run>.dat", "indentationShowTipSurf": "yes", "indentationSurfaceSize": size,
        "indentationSurfaceStep": grid_step, "indentationCantLength": cant_len,
        "indentationSurfConnectFile": "mica.vmd", "outputname": "<name>_<run>.dat",
        "finalcoord": "<name>_<run>_final.pdb", "reffilename": "<name>.ref.pdb",
    }
    write_cfg("run.sop", cfg)
    print("case: indentation, N=40, run %d, %d steps" % (run, steps), flush=True)
    S.run_config("run.sop")
    p = O.make_params(pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=0.16, temperature=0.6,
                      seed=seed + run, nav=100, outputtiming=100, dcdfreq=400, indentation=1,
                      indChip=tip, indTip=tip, indDirection=(0.0, 0.0, -1.0),
                      indSurfR0=(0.0, 0.0, surf_z), indSurfN=(0.0, 0.0, 1.0), indTipR=8.0,
                      indKs=0.07, indDeltaX=0.8, indTipZeta=500.0, indOutputFreq=200)
    ref = O.Sim(p, sys_)
    ref.run(steps)
    n = sys_.n
    final = atom_lines("tiny_%d_final.pdb" % run)
    assert [line[30:54] for line in final[:n]] == oracle_xyz_text(ref.coords())
    rows = dat_rows("tiny_%d.dat" % run)
    want = ref.dat()
    assert len(rows) == len(want) == steps // 100 + 1
    for got, w in zip(rows, want):
        assert got[:7] == oracle_dat_text(w)
    with open("ind_%d.dat" % run) as f:
        got = [line.rstrip("\n") for line in f]
    want = [ind_text(r) for r in ref.ind_rows()]
    assert len(want) == steps // 200 + 1
    assert got == want
    # representation beads (indentation.cu:254-300): the cantilever base
    # (chip + a vector of the cantilever's length perpendicular to the
    # direction), the tip, and a size x size grid in the surface plane
    ref_atoms = atom_lines("tiny.ref.pdb")
    assert len(ref_atoms) == n + 2 + size * size
    extra = [tuple(float(line[30 + 8 * k:38 + 8 * k]) for k in range(3)) for line in ref_atoms[n:]]
    base = np.array(tip) + np.array([1.0, 0.0, 0.0]) * cant_len  # direction (0,0,-1): r = (1, 0, 0)
    np.testing.assert_allclose(extra[0], base, atol=1e-3)
    np.testing.assert_allclose(extra[1], tip, atol=1e-3)
    nrm = np.array([0.0, 0.0, 1.0])
    r1 = np.array([1.0, 0.0, 0.0])  # the first axis with a zero normal component
    r2 = np.cross(nrm, r1)
    for i in range(size):
        for j in range(size):
            q = np.array([0.0, 0.0, surf_z]) + (i - size // 2) * grid_step * r1 + (j - size // 2) * grid_step * r2
            np.testing.assert_allclose(extra[2 + i * size + j], q, atol=1e-3)
    assert [line[17:20] for line in ref_atoms[n:n + 2]] == ["TIP", "TIP"]
    assert all(line[17:20] == "MIC" for line in ref_atoms[n + 2:])
    # VMD script: the two cantilever beads bonded to each other, each surface
    # bead to its +i and +j neighbours
    want = []
    for a, b in ((n, n + 1), (n + 1, n)):
        want += ['set sel [atomselect top "index %d"]' % a, "set bonds {%d}" % b, "$sel setbonds [list $bonds]",
                 "$sel delete"]
    for i in range(size):
        for j in range(size):
            k = n + 2 + i * size + j
            bonds = ("%d" % (k + size) if i != size - 1 else "") + (" %d" % (k + 1) if j != size - 1 else "")
            want += ['set sel [atomselect top "index %d"]' % k, "set bonds {%s}" % bonds,
                     "$sel setbonds [list $bonds]", "$sel delete"]
    with open("mica.vmd") as f:
        assert f.read().splitlines() == want
