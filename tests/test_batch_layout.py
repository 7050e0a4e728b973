"""Byte-level audit of the batched child-LP kernel's memory layout, on the
host: tests/batch_layout_audit.cpp, compiled with g++ against the layout
header the kernel and its host code use (sovereign_amd/csrc/
batch_layout.hpp), walks every workspace region and every 16-byte load of
the FP32 assembly and every 32-byte load of the FP64 assembly for all row
counts 1..272 against the allocation ends. Three planted mistakes (a
workspace 16 bytes short, the FP32 copy of A one float short, the FP64 copy
one double short) must make it fail."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def audit(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("audit") / "batch_layout_audit")
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "sovereign_amd", "csrc"),
                    "-o", exe, os.path.join(HERE, "batch_layout_audit.cpp")], check=True)
    return exe


def test_batch_layout_in_bounds(audit):
    r = subprocess.run([audit], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok ")


@pytest.mark.parametrize("plant", ["w", "a", "d"])
def test_batch_layout_audit_catches_planted_mistakes(audit, plant):
    r = subprocess.run([audit, plant], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "FAIL" in r.stdout
