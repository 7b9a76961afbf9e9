"""File formats and the RNG seed table against the golden fixtures made by
tests/golden/make_golden.py from the reference's own host code
(src/Integrators/ran2.cpp, src/IO/topio.cpp, src/IO/pdbio.cpp,
src/IO/dcdio.cpp).  CPU only: the readers and writers of the library and
the oracle's seed table."""
import json
import os

import numpy as np
import pytest

import oracle as O
import sopgpu as S

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return os.path.join(G, name)


@pytest.fixture(scope="module")
def rng_golden():
    with open(golden("rng.json")) as f:
        return json.load(f)


def test_oracle_seed_table_matches_reference(rng_golden):
    for case in rng_golden["seeds"]:
        want = np.array(case["table"], dtype=np.uint32)
        np.testing.assert_array_equal(O.generate_seeds(case["seed"], case["n"]), want, err_msg=str(case["seed"]))


def test_library_seed_table_matches_reference(rng_golden):
    for case in rng_golden["seeds"]:
        want = np.array(case["table"], dtype=np.uint32)
        np.testing.assert_array_equal(S.generate_seeds(case["seed"], case["n"]), want, err_msg=str(case["seed"]))


def test_seed_table_rows_are_ran2_draws(rng_golden):
    # ht.cu:40-56: every word is (unsigned)(ran2 * UINT_MAX), redrawn while < 128;
    # a negative seed initialises ran2's shuffle table, so seed -12345's table
    # starts with the stream of ran2(-12345) scaled to 32 bits.
    r = next(c for c in rng_golden["ran2"] if c["idum"] == -12345)
    vals = [float(v) for v in r["values"]]
    words = [int(v * 4294967295.0) for v in vals]
    words = [w for w in words if w >= 128]
    table = S.generate_seeds(-12345, 8).reshape(-1)
    assert list(table[:len(words[:32])]) == words[:32]


def _dump_rows(path, tag):
    with open(path) as f:
        return [line.split()[1:] for line in f if line.split()[0] == tag]


def test_read_topology_matches_reference_parse():
    s = S.read_topology(golden("tiny.top"))
    rows_b = _dump_rows(golden("tiny_top.txt"), "B")
    rows_n = _dump_rows(golden("tiny_top.txt"), "N")
    rows_p = _dump_rows(golden("tiny_top.txt"), "P")
    rows_a = _dump_rows(golden("tiny_top.txt"), "A")
    assert s.n == len(rows_a) == 40
    np.testing.assert_array_equal(s.bonds, np.array([[int(r[0]), int(r[1])] for r in rows_b]))
    np.testing.assert_array_equal(s.bond_r0, np.array([float(r[2]) for r in rows_b], np.float32))
    np.testing.assert_array_equal(s.natives, np.array([[int(r[0]), int(r[1])] for r in rows_n]))
    np.testing.assert_array_equal(s.nat_r0, np.array([float(r[2]) for r in rows_n], np.float32))
    np.testing.assert_array_equal(s.nat_eh, np.array([float(r[3]) for r in rows_n], np.float32))
    np.testing.assert_array_equal(s.pairs, np.array([[int(r[0]), int(r[1])] for r in rows_p]))


@pytest.mark.parametrize("src,dump", [("tiny.pdb", "tiny_coords_pdb.txt"), ("tiny.xyz", "tiny_coords_xyz.txt")])
def test_read_coordinates_matches_reference(src, dump):
    s = S.read_topology(golden("tiny.top"), golden(src))
    want = np.loadtxt(golden(dump), dtype=np.float64).astype(np.float32)
    np.testing.assert_array_equal(s.coords[:, :3], want)
    assert not s.coords[:, 3].any()


def test_write_pdb_is_byte_identical_to_reference(tmp_path):
    s = S.read_topology(golden("tiny.top"), golden("tiny.pdb"))
    out = str(tmp_path / "saved.pdb")
    S.write_pdb(golden("tiny.top"), s.coords, out)
    with open(out, "rb") as a, open(golden("tiny_saved.pdb"), "rb") as b:
        assert a.read() == b.read()


def test_write_dcd_matches_reference_except_date(tmp_path):
    n, frames = 7, 3
    f = np.zeros((frames, 3, n), np.float32)
    for k in range(frames):
        i = np.arange(n, dtype=np.float32)
        f[k, 0] = np.float32(0.5) * i + np.float32(k)
        f[k, 1] = np.float32(-0.25) * i + np.float32(2.0) * np.float32(k)
        f[k, 2] = np.float32(1.0) / (np.float32(1.0) + i) - np.float32(k)
    out = str(tmp_path / "t.dcd")
    S.write_dcd(out, f, freq=1000)
    a = open(out, "rb").read()
    b = open(golden("tiny.dcd"), "rb").read()
    assert len(a) == len(b)
    date = slice(4 * 23 + 8 + 80, 4 * 23 + 8 + 160)  # the second title record
    assert a[:date.start] == b[:date.start] and a[date.stop:] == b[date.stop:]
    assert a[date][:14] == b[date][:14] == b"REMARKS DATE: "


def test_malformed_topology_is_reported(tmp_path):
    p = tmp_path / "bad.top"
    p.write_text("[ atoms ]\n0 CA 1 ALA CA A 1.0 0.0\n[ angles ]\n")
    with pytest.raises(S.SopError, match="Malformed section header"):
        S.read_topology(str(p))
    p.write_text("0 1 1 3.8\n")
    with pytest.raises(S.SopError, match="outside any section"):
        S.read_topology(str(p))


def test_coordinate_count_mismatch_is_reported(tmp_path):
    p = tmp_path / "short.pdb"
    lines = open(golden("tiny.pdb")).read().splitlines()
    p.write_text("\n".join(lines[:10]) + "\n")
    with pytest.raises(S.SopError, match="Read coordinates for 9 beads, yet topology has 40 beads"):
        S.read_topology(golden("tiny.top"), str(p))
