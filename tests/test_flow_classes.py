"""Host-side flow structure (csrc/qap_pt_flow.cpp) against brute force, on the CPU.

The sweep kernels skip the force update for swaps inside a facility class and take a
two-column update for swaps of twins; both are only correct if the classes the host
computes are exact.  tests/native/flow_classes_check.cpp is compiled with g++ from that
one source file and compares every pair of facilities with the definitions.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..",
                    "p-dit-probabilistic-ising-machine-for-solving-the-quadratic-assignment-problem_amd",
                    "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_flow_classes_match_brute_force(tmp_path):
    exe = tmp_path / "flow_classes_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", str(exe),
                    os.path.join(HERE, "native", "flow_classes_check.cpp"),
                    os.path.join(CSRC, "qap_pt_flow.cpp")], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
