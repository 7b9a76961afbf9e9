"""Writes the golden fixtures of this directory.

Runs in a checkout where the reference's host sources are present, through
oracle/_ref/ref_harness (built by oracle/Makefile from the reference's own
src/Integrators/ran2.cpp, src/IO/topio.cpp, src/IO/pdbio.cpp,
src/IO/dcdio.cpp).  The outputs are data: seed tables and ran2 streams,
the reference's parse of a small synthetic topology/coordinate set, and the
PDB and DCD files its writers produce.

    python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

SEED_CASES = [(1234, 64), (4242, 300), (-7, 16), (0, 8), (2147483000, 8)]
RAN2_CASES = [(-12345, 200), (17, 50), (0, 20)]


def run(*args):
    return subprocess.run([HARNESS, *map(str, args)], check=True, capture_output=True, text=True).stdout


def tagged(out, tags):
    return [line for line in out.splitlines() if line.split(" ", 1)[0] in tags]


def write_topology(path_top, path_pdb, path_xyz):
    """A two-chain system of 40 beads written in the .top layout, with
    comments, blank lines, irregular spacing and a [ pairs ] section."""
    rng = np.random.default_rng(20261202)
    n0, n1 = 24, 16
    xyz = []
    for c, n in enumerate((n0, n1)):
        p = np.array([10.0 * c, -5.0 * c, 2.0 * c])
        for i in range(n):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            p = p + 3.8 * d
            xyz.append(p.copy())
    xyz = np.array(xyz)
    n = n0 + n1
    with open(path_top, "w") as f:
        f.write("; synthetic topology for reader tests\n\n[ atoms ]\n")
        f.write(";   nr       type  resnr residue  atom   cgnr     charge       mass\n")
        for i in range(n):
            ch = "A" if i < n0 else "B"
            res = ["ALA", "GLY", "TRP", "LYS"][i % 4]
            f.write("%6d%11s%7d%7s%7s%7s%11.2f%11.3f\n" % (i, "CA", i % n0 + 1, res, "CA", ch, 1.0, 0.25 * (i % 3)))
        f.write("\n[ bonds ]\n;  ai    aj funct            c0\n")
        for i in range(n - 1):
            if (i < n0) == (i + 1 < n0):
                r0 = np.linalg.norm(xyz[i + 1] - xyz[i])
                sep = "  " if i % 2 else " "
                f.write("%5d%s%5d %5d %10.5f\n" % (i, sep, i + 1, 1, r0))
        f.write("\n[ native ]\n;  ai    aj funct            c0            c1\n")
        for i in range(n):
            for j in range(i + 3, n):
                r = np.linalg.norm(xyz[j] - xyz[i])
                if r < 9.0:
                    f.write("%5d %5d %5d %10.5f %10.5f\n" % (i, j, 1, r, 0.9 + 0.6 * rng.random()))
        f.write("\n[ pairs ]\n")
        for i in range(0, n, 5):
            for j in range(i + 7, n, 9):
                f.write("%5d %5d %5d\n" % (i, j, 1))
    with open(path_pdb, "w") as f:
        f.write("REMARK synthetic coordinates\n")
        for i in range(n):
            f.write("ATOM  %5d  CA  ALA %s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n"
                    % (i + 1, "A" if i < n0 else "B", i + 1, *xyz[i], 1.0, 0.0))
        f.write("END\n")
    with open(path_xyz, "w") as f:
        f.write("%d\nsynthetic\n" % n)
        for i in range(n):
            f.write("CA\t%.4f %.4f\t%.4f\n" % tuple(xyz[i]))


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness is not built (reference sources absent)")
    rng_golden = {"seeds": [], "ran2": []}
    for seed, n in SEED_CASES:
        rows = [[int(v) for v in line.split()] for line in run("seeds", seed, n).strip().splitlines()]
        rng_golden["seeds"].append({"seed": seed, "n": n, "table": rows})
    for idum, count in RAN2_CASES:
        rows = [line.split() for line in run("ran2", idum, count).strip().splitlines()]
        rng_golden["ran2"].append({"idum": idum, "count": count, "values": [r[0] for r in rows],
                                   "idum_after": [int(r[1]) for r in rows]})
    with open(os.path.join(HERE, "rng.json"), "w") as f:
        json.dump(rng_golden, f, indent=0)

    top = os.path.join(HERE, "tiny.top")
    pdb = os.path.join(HERE, "tiny.pdb")
    xyz = os.path.join(HERE, "tiny.xyz")
    write_topology(top, pdb, xyz)
    with open(os.path.join(HERE, "tiny_top.txt"), "w") as f:
        f.write("\n".join(tagged(run("top", top), {"A", "B", "N", "P", "atoms", "bonds", "natives", "pairs"})) + "\n")
    for src, name in ((pdb, "tiny_coords_pdb.txt"), (xyz, "tiny_coords_xyz.txt")):
        out = [line for line in run("coords", top, src).splitlines() if len(line.split()) == 3
               and all(c in "0123456789.-+e" for c in line.replace(" ", ""))]
        with open(os.path.join(HERE, name), "w") as f:
            f.write("\n".join(out) + "\n")
    run("savepdb", top, pdb, os.path.join(HERE, "tiny_saved.pdb"))
    run("dcd", os.path.join(HERE, "tiny.dcd"), 7, 3)


if __name__ == "__main__":
    main()
