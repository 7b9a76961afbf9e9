"""The sop-gpu front end's parameter handling on the CPU: the errors the
reference raises for its parameter files (src/sop-gpu.cpp:18-88,
src/gsop.cu:66-70, src/Potentials/pulling.cu:47-63,96-117,
src/Integrators/bdhitea.cu:28-33), reported before any device is used."""
import os
import subprocess

import pytest

import sopgpu as S
from conftest import gpu_present

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sop-gpu_amd", "sop-gpu")

BASE = {
    "name": "tiny",
    "topology": os.path.join(GOLDEN, "tiny.top"),
    "coordinates": os.path.join(GOLDEN, "tiny.pdb"),
    "runnum": "2",
    "firstrun": "0",
    "numsteps": "100",
    "timestep": "0.16",
}


def write_cfg(tmp_path, drop=(), **extra):
    kv = {k: v for k, v in BASE.items() if k not in drop}
    kv.update({k: str(v) for k, v in extra.items()})
    p = tmp_path / "run.sop"
    lines = ["# parameter file", ""]
    for k, v in kv.items():
        lines.append("%s\t%s\t# comment" % (k, v))
    p.write_text("\n".join(lines) + "\n")
    return str(p)


@pytest.mark.parametrize("missing", ["numsteps", "topology", "runnum"])
def test_mandatory_parameter_missing(tmp_path, missing):
    with pytest.raises(S.SopError, match="Parameter '%s' not specified" % missing):
        S.run_config(write_cfg(tmp_path, drop=(missing,)))


def test_stage_parameter_rejected(tmp_path):
    with pytest.raises(S.SopError, match="`stage` is no longer supported"):
        S.run_config(write_cfg(tmp_path, stage="pull"))


def test_missing_topology_file(tmp_path):
    with pytest.raises(S.SopError, match="Can't open topology file"):
        S.run_config(write_cfg(tmp_path, topology=str(tmp_path / "none.top")))


def test_pulling_needs_deltax_or_fconst(tmp_path):
    cfg = write_cfg(tmp_path, pulling="on", fixed="0", pulled="39", fixedEnd=0, pulledEnd=39)
    with pytest.raises(S.SopError, match="Either 'deltax' or 'fconst'"):
        S.run_config(cfg)


def test_pulling_deprecated_bead_lists(tmp_path):
    cfg = write_cfg(tmp_path, pulling="on", deltax=0.01, fixed_beads="0", pulled="39", fixedEnd=0, pulledEnd=39)
    with pytest.raises(S.SopError, match="deprecated"):
        S.run_config(cfg)


def test_pulling_bead_out_of_range(tmp_path):
    cfg = write_cfg(tmp_path, pulling="on", deltax=0.01, fixed="0", pulled="41", fixedEnd=0, pulledEnd=39)
    with pytest.raises(S.SopError, match="Invalid pulling bead #41"):
        S.run_config(cfg)


def test_bad_pull_direction(tmp_path):
    cfg = write_cfg(tmp_path, pulling="on", deltax=0.01, fixed="0", pulled="39", fixedEnd=0, pulledEnd=39,
                    pullDirection="sideways")
    with pytest.raises(S.SopError, match="'endToEnd' or 'vector'"):
        S.run_config(cfg)


def test_minimization_with_hi_rejected(tmp_path):
    cfg = write_cfg(tmp_path, minimization="yes", hi_on="yes", hi_epsilon_freq=10)
    with pytest.raises(S.SopError, match="minimization with BDHI"):
        S.run_config(cfg)


def test_bad_boolean_and_float3(tmp_path):
    with pytest.raises(S.SopError, match="boolean parameter 'pulling'"):
        S.run_config(write_cfg(tmp_path, pulling="maybe"))
    cfg = write_cfg(tmp_path, indentation="on", indentationChip="1 2", indentationDirection="0 0 1",
                    indentationSurfaceR0="0 0 0", indentationSurfaceN="0 0 1", indentationTipR=10,
                    indentationKs=0.05, indentationDeltaX=0.01, runnum=1)
    with pytest.raises(S.SopError, match="Float3 parameter 'indentationChip' has 2 components"):
        S.run_config(cfg)


def test_indentation_many_runs_rejected(tmp_path):
    cfg = write_cfg(tmp_path, indentation="on", indentationChip="0 0 60", indentationDirection="0 0 -1",
                    indentationSurfaceR0="0 0 -60", indentationSurfaceN="0 0 1", indentationTipR=10,
                    indentationKs=0.05, indentationDeltaX=0.01)
    with pytest.raises(S.SopError, match="Many-runs-per-GPU mode is not supported by indentation"):
        S.run_config(cfg)


def test_masks_and_array_ranges(tmp_path):
    # `<name>` masks and `a to b` ranges resolve before the first file is read:
    # the coordinates path below only exists once the mask is applied.
    cfg = write_cfg(tmp_path, coordinates=os.path.join(GOLDEN, "<name>.pdb"), pulling="on", deltax=0.01,
                    fixed="0 to 3 7", pulled="39", fixedEnd=0, pulledEnd=39)
    if gpu_present():
        pytest.skip("on a GPU the run proceeds; covered by the GPU config tests")
    with pytest.raises(S.SopError, match="HIP device"):
        S.run_config(cfg)


def test_cli_reports_error_and_exit_code(tmp_path):
    r = subprocess.run([CLI, write_cfg(tmp_path, drop=("numsteps",))], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Parameter 'numsteps' not specified" in r.stderr
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Configuration file should be specified" in r.stderr
