"""The host thread pool that splits a large step's staging copies
(gpusharesat_amd/csrc/stage_pool.h), built with g++ and ThreadSanitizer and
run on the CPU: every part of every job runs once, jobs never overlap, an
early return waits for the running job, destruction joins the workers."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _build(tmp_path, flags):
    exe = tmp_path / "stage_pool_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", *flags,
           "-I", str(ROOT / "gpusharesat_amd" / "csrc"),
           str(ROOT / "tests" / "cpp" / "stage_pool_check.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return exe if r.returncode == 0 else None


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_stage_pool_under_thread_sanitizer(tmp_path):
    exe = _build(tmp_path, ["-fsanitize=thread"])
    if exe is None:  # no TSan runtime: the plain build still checks the logic
        exe = _build(tmp_path, [])
    assert exe is not None, "stage_pool_check did not build"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
