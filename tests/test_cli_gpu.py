"""The command-line drivers: they run, and a run that is snapshotted (-s) and
resumed (-r) ends exactly where the uninterrupted run ends (MPM and DEM on the
shipped column-collapse scenes: same printed state to the last digit; hybrid
on a packed grain column with the shipped coupling settings and a resampling
pass every 4 steps: the same reported state, across passes, and the passes
fall on the reference's schedule).
"""
import subprocess
from pathlib import Path

import pytest

from ._util import write_packed_hybrid_scene

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "libhybridgrains_amd" / "bin"
SCENES = ROOT / "tests" / "golden" / "assets" / "column-collapse"


def _run(cwd, *args):
    r = subprocess.run([str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _last_state(out):
    return [ln for ln in out.splitlines() if ln.startswith("step ")][-1]


@pytest.mark.parametrize("cli,scene", [("mpmgrains2d_cli", "mpm.xml"),
                                       ("rigidbody2d_cli", "dem.xml")])
def test_snapshot_resume_matches_straight_run(tmp_path, cli, scene):
    exe, xml = BIN / cli, SCENES / scene
    straight = _last_state(_run(tmp_path, exe, "-n", "30", xml))
    _run(tmp_path, exe, "-n", "12", "-s", "0", xml)
    assert (tmp_path / "serial.bin").stat().st_size > 0
    resumed = _last_state(_run(tmp_path, exe, "-n", "30", "-r", "serial.bin", xml))
    assert resumed.startswith("step 30 ")
    assert resumed == straight


def _state_lines(out):
    return {int(ln.split()[1]): ln for ln in out.splitlines() if ln.startswith("step ")}


def test_hybrid_cli_resamples_on_the_reference_schedule(tmp_path):
    # avoidAVoidFreq 3: two passes at start-up (initializeAvoidAVoid), then the
    # step counter runs 1 -> 3 through the first step's two opening calls and
    # a pass ends every 4th step: steps 2, 6, 10
    xml = write_packed_hybrid_scene(tmp_path / "scene", freq=3)
    out = _run(tmp_path, BIN / "hybridgrains2d_cli", "-n", "10", "-f", "1", xml)
    lines = _state_lines(out)
    passes = {s: int(ln.split()[ln.split().index("resamplings") + 1]) for s, ln in lines.items()}
    assert passes == {1: 2, 2: 3, 3: 3, 4: 3, 5: 3, 6: 4, 7: 4, 8: 4, 9: 4, 10: 5}
    last = lines[10].split()
    grains, points = int(last[last.index("grains") + 1]), int(last[last.index("points") + 1])
    # the continuum was seeded inside the packing, whose grains there were removed
    assert points > 0 and grains < 1360


def test_hybrid_cli_snapshots_and_resumes(tmp_path):
    exe = BIN / "hybridgrains2d_cli"
    xml = write_packed_hybrid_scene(tmp_path / "scene", freq=3)
    straight = _state_lines(_run(tmp_path, exe, "-n", "10", "-f", "2", xml))
    out = _run(tmp_path, exe, "-n", "6", "-f", "3", "-s", "1", xml)
    assert "Serialized step 6 to serial_0000000006.bin" in out
    assert (tmp_path / "serial_0000000003.bin").exists()
    out = _run(tmp_path, exe, "-n", "10", "-f", "2", "-r", "serial_0000000006.bin", xml)
    assert "Simulation complete: 4 steps" in out
    resumed = _state_lines(out)
    # the resumed run crosses the pass at step 10 as the straight run does
    assert resumed[8] == straight[8] and resumed[10] == straight[10]
    r = subprocess.run([str(BIN / "mpmgrains2d_cli"), "-n", "3", "-r",
                        "serial_0000000006.bin", str(SCENES / "mpm.xml")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "does not appear to be a snapshot" in r.stderr
