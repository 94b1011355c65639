"""Runs the C++ host-surface checks (tests/cpp/test_host.cpp, built into pilot_amd/lib)."""

import os
import subprocess

import pytest

from pilot_amd import PACKAGE_DIR

EXE = os.path.join(PACKAGE_DIR, "lib", "test_host")


def _run(mode):
    assert os.path.exists(EXE), "build the host surface first (make -C pilot_amd/csrc host)"
    p = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr


def test_host_surface_cpu():
    _run("cpu")


@pytest.mark.gpu
def test_host_surface_gpu():
    _run("gpu")
