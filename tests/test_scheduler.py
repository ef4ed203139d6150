"""The scene server's scheduling policy (libhybridgrains_amd/host/hg_sched.h)
against a host-only executor on a simulated clock: admission by memory
budget, bounded bypass, cancellation and failure isolation, latency
quantiles, continuous admission.  Compiled with g++ and run here."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_scheduler_policy(tmp_path):
    exe = tmp_path / "test_sched"
    subprocess.run(
        ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
         f"-I{ROOT / 'libhybridgrains_amd' / 'host'}",
         str(ROOT / "tests" / "cpp" / "test_sched.cpp"), "-o", str(exe)],
        check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all scheduler checks passed" in r.stdout
