"""The command-line drivers on the shipped column-collapse scenes: they run,
and a run that is snapshotted (-s) and resumed (-r) ends exactly where the
uninterrupted run ends (MPM and DEM: same printed state to the last digit;
hybrid: it resumes and completes).
"""
import subprocess
from pathlib import Path

import pytest

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


def test_hybrid_cli_snapshots_and_resumes(tmp_path):
    exe, xml = BIN / "hybridgrains2d_cli", SCENES / "hybrid.xml"
    out = _run(tmp_path, exe, "-n", "6", "-f", "3", "-s", "1", xml)
    assert "Serialized step 6 to serial_0000000006.bin" in out
    assert (tmp_path / "serial_0000000003.bin").exists()
    out = _run(tmp_path, exe, "-n", "10", "-f", "2", "-r", "serial_0000000006.bin", xml)
    assert "step 8 " in out and "step 10 " in out and "Simulation complete: 4 steps" in out
    r = subprocess.run([str(BIN / "mpmgrains2d_cli"), "-n", "3", "-r",
                        "serial_0000000006.bin", str(SCENES / "mpm.xml")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "does not appear to be a snapshot" in r.stderr
