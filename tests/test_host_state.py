"""Builds and runs tests/cpp/test_host_state.cpp (CPU only): the sharer's host
state — assignment windows, clause activity policy, per-solver delivery —
against the reference's known answers."""
import os
import subprocess

TESTS = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(TESTS)


def test_host_state_known_answers(tmp_path):
    exe = tmp_path / "test_host_state"
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "gpusharesat_amd", "csrc"),
           os.path.join(TESTS, "cpp", "test_host_state.cpp"),
           os.path.join(ROOT, "gpusharesat_amd", "csrc", "host_state.cpp"), "-o", str(exe),
           "-pthread"]
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
