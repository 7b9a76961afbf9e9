"""Host plugins through the C ABI (sop_sim_*, include/sop_plugin.h), the
counterpart of the reference's SOPPotential / SOPUpdater registries
(src/gsop.h:69-105, src/gsop.cu:31-34,80-101,237-253).

On the CPU: the C++ header compiles and links against the library, and the
error paths (a missing parameter file, a null simulation, an unknown name)
report the reference's messages.  On the GPU: host updaters run every
stride steps after the built-in ones in registration order, their
frequencies shorten the stride as any updater's does, a host potential's
forces enter the step (a free bead drifts by exactly (h/zeta) F per step in
FP32), its energy term is a column of the energy .dat, and a failing plugin
ends the run with its message."""
import os
import subprocess

import numpy as np
import pytest

import sopgpu as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PLUGIN_CPP = r'''
#include "sop_plugin.h"
#include <cstdio>
#include <cstring>

static int released = 0;
struct Counter : sop::SOPUpdater {
    Counter() { name = "Counter"; frequency = 20; }
    ~Counter() override { ++released; }
    void update() override {}
};
struct Push : sop::SOPPotential {
    Push() { name = "Push"; }
    void compute() override {}
    int getEnergiesCount() const override { return 1; }
    float getEnergy(int, int) override { return 0.5f; }
};
int main(int argc, char** argv) {
    sop_sim* sim = nullptr;
    if (sop_sim_create(argv[1], 0, nullptr, nullptr, &sim) != SOP_E_CONFIG || sim) return 1;
    if (!std::strstr(sop_last_error(), "missing.sop")) return 2;
    // a failed registration still hands the object over (and releases it)
    if (sop::addUpdater(nullptr, new Counter) != SOP_E_ARG || released != 1) return 3;
    if (sop::addPotential(nullptr, new Push) != SOP_E_ARG) return 4;
    if (sop::potentialByName(nullptr, "Covalent") != -1) return 5;
    std::printf("ok\n");
    return 0;
}
'''


def test_plugin_header_compiles_links_and_reports_errors(tmp_path):
    src = tmp_path / "plug.cpp"
    src.write_text(PLUGIN_CPP)
    exe = tmp_path / "plug"
    lib_dir = os.path.join(ROOT, "sop-gpu_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", lib_dir, "-lsopgpu", "-Wl,-rpath," + lib_dir], check=True)
    out = subprocess.run([str(exe), str(tmp_path / "missing.sop")], capture_output=True, text=True)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    assert out.stdout.strip().endswith("ok")


def test_simulation_errors_without_a_gpu(tmp_path):
    with pytest.raises(S.SopError) as e:
        S.Simulation(str(tmp_path / "missing.sop"))
    assert e.value.code == -6
    assert S.lib.sop_sim_potential_by_name(None, b"Covalent", None) == -1
    assert S.lib.sop_sim_updater_name(None, 0) is None


def _free_bead_files(d, xyz):
    """A topology of beads with no bonds, contacts or pairs, and its PDB."""
    top = ["[ atoms ]"]
    pdb = ["REMARK free beads"]
    for k, (x, y, z) in enumerate(xyz):
        top.append("%6d         CA %6d    ALA     CA      A       1.00      0.000" % (k, k + 1))
        pdb.append("ATOM  %5d  CA  ALA A%4d    %8.3f%8.3f%8.3f  1.00  0.00" % (k + 1, k + 1, x, y, z))
    top += ["[ bonds ]", "[ native ]"]
    with open(os.path.join(d, "free.top"), "w") as f:
        f.write("\n".join(top) + "\n")
    with open(os.path.join(d, "free.pdb"), "w") as f:
        f.write("\n".join(pdb + ["END"]) + "\n")


def _cfg(d, **kw):
    cfg = {"name": "free", "topology": "free.top", "coordinates": "free.pdb", "run": 0, "numsteps": 390,
           "timestep": 0.16, "zeta": 50.0, "temperature": 0.0, "seed": 11, "nav": 100, "outputtiming": 100,
           "dcdfreq": 100, "pairs_freq": 100, "possiblepairs_freq": 100, "pairs_threshold": 32.0,
           "outputname": "<name>.dat", "finalcoord": "<name>_final.pdb"}
    cfg.update(kw)
    with open(os.path.join(d, "run.sop"), "w") as f:
        for k, v in cfg.items():
            f.write("%s\t%s\n" % (k, v))
    return os.path.join(d, "run.sop")


@pytest.mark.gpu
def test_host_plugins_run_in_the_loop(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    xyz = [(0.0, 0.0, 0.0), (40.0, 0.0, 0.0), (0.0, 40.0, 0.0)]
    _free_bead_files(str(tmp_path), xyz)
    calls = []
    force = np.zeros((3, 4), np.float32)
    force[1, :3] = (0.3, -0.2, 0.1)
    with S.Simulation(_cfg(str(tmp_path))) as sim:
        assert sim.potentials() == ["Covalent", "Native", "Long Range"]
        assert [u[0] for u in sim.updaters()] == ["Possiblepairlist", "DAT output", "Pairlist", "DCD output"]
        assert sim.integrator() == "Langevin"
        sim.add_updater("A", 40, lambda: calls.append(("A", sim.state()[0])))
        sim.add_updater("B", 25, lambda: calls.append(("B", sim.state()[0])))
        sim.add_potential("Push", lambda: sim.add_forces(0, force), 1, lambda traj, i: 0.25)
        assert sim.potential_by_name("Push") == 3 and sim.updater_by_name("B") == 5
        with pytest.raises(S.SopError) as e:
            sim.potential_by_name("Gravity")
        assert "Unable to find requested potential 'Gravity'" in str(e.value)
        print("case: host plugins, 3 free beads, 400 steps", flush=True)
        sim.run()
        final = sim.coords(0)
        with pytest.raises(S.SopError):
            sim.run()  # once only
    # every pass calls A then B; the passes are every 25 steps (the stride
    # min(nav, 100, 40, 25)), through step 375 (numsteps 390)
    passes = list(range(0, 400, 25))
    assert calls == [(w, s) for s in passes for w in ("A", "B")]
    # bead 1 drifted by (h/zeta) F per step, in FP32, the others not at all
    m = np.float32(np.float32(0.16) / np.float32(50.0))
    x = np.array(xyz[1], np.float32)
    for _ in range(400):
        x = (x + (m * force[1, :3]).astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(final[1, :3], x)
    np.testing.assert_array_equal(final[[0, 2], :3], np.array([xyz[0], xyz[2]], np.float32))
    # the host potential's term is the energy .dat's fourth energy column
    with open("free.dat") as f:
        rows = [line.rstrip("\n").split("\t") for line in f]
    assert [int(r[0]) for r in rows] == [0, 100, 200, 300]
    assert all(r[5] == "%8.5f" % 0.25 and r[6] == "%8.5f" % 0.25 for r in rows)


@pytest.mark.gpu
def test_failing_plugin_ends_the_run_with_its_message(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    _free_bead_files(str(tmp_path), [(0.0, 0.0, 0.0), (40.0, 0.0, 0.0)])

    def update():
        if sim.state()[0] >= 50:
            raise RuntimeError("bead out of the box at step %d" % sim.state()[0])

    sim = S.Simulation(_cfg(str(tmp_path)))
    try:
        sim.add_updater("Guard", 50, update)
        with pytest.raises(S.SopError) as e:
            sim.run()
        assert e.value.code == -6 and "bead out of the box at step 50" in str(e.value)
    finally:
        sim.close()


def test_integration_example_compiles(tmp_path):
    """The custom updater / potential example of INTEGRATION.md builds
    against include/sop_plugin.h and links with the library."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index("## 2. Custom potentials")
    block = text[text.index("```cpp", start) + len("```cpp"):]
    block = block[:block.index("```")]
    src = tmp_path / "example.cpp"
    src.write_text(block)
    lib_dir = os.path.join(ROOT, "sop-gpu_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(tmp_path / "example"), "-L", lib_dir, "-lsopgpu", "-Wl,-rpath," + lib_dir], check=True)
