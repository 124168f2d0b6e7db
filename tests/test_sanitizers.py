"""CPU sanitizer builds: the oracle over small diagrams of every kind the
tests use, and the host runtime's own logic (shell-table builder, grid size,
argument checks), each under AddressSanitizer + UndefinedBehaviorSanitizer.
Any report, or a wrong result, fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args):
    r = subprocess.run(args, cwd=ROOT, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    return out


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_asan_ubsan():
    out = _run(["make", "-s", "-C", "oracle", "sanitize"])
    assert "sanitized oracle run: clean" in out


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="needs hipcc")
def test_host_runtime_under_asan_ubsan():
    # the device objects come from the normal build
    _run(["make", "-s", "-j8", "-C", "the-tessellator_amd", "all"])
    out = _run(["make", "-s", "-j8", "-C", "the-tessellator_amd", "sanitize"])
    assert "sanitized host run: clean" in out
