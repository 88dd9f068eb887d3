"""Statistical parity with the reference's own CPU solver -- CPU only.

oracle/_ref/cpu_pim_ref is CPU-PIM/CPU_PIM.cpp compiled unmodified (oracle/Makefile,
target `ref`; needs the reference checkout, so the test skips where it is absent).
Its RNG is an unseeded mt19937 (CPU_PIM.cpp:110,130), so single runs cannot be compared
bit for bit; the comparison is between distributions of the best cost after a fixed
number of exchange intervals, reference vs the oracle's reference-faithful
floating-point mode and its exact integer mode (the definition the GPU reproduces).
The reference folds replica bests into its reported value at the start of each
mega-iteration (CPU_PIM.cpp:297-321), so `-w W` reports the best after W-1 intervals.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle

pytestmark = pytest.mark.skipif(not os.path.exists(pyoracle.REF_BIN),
                                reason="reference binary not built (needs the reference checkout)")

N, R, T, SWEEPS, RUNS = 40, 10, 4, 12, 24
BETA0_LN, RATIO = 0.02, 1.5


def _ref_best(path, outdir):
    cmd = [pyoracle.REF_BIN, "-b", str(BETA0_LN), "-m", str(RATIO), "-r", str(R), "-t", str(T),
           "-w", str(SWEEPS + 1), "-i", "1", "-o", outdir, path]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=60, check=True).stdout
    return int(re.search(r"Single best energy found: (-?\d+)", out).group(1))


def test_best_cost_distribution_matches_reference(qpt, oracle):
    F, D = qpt.instances.mixed_instance(77, N, dense=True)
    betas = qpt.instances.beta_ladder_from_flags(R, BETA0_LN, RATIO)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "inst.dat")
        qpt.instances.write_qaplib(path, F, D)
        ref = np.array([_ref_best(path, os.path.join(tmp, f"o{k}")) for k in range(RUNS)], float)
    fp = np.array([oracle.solve(F, D, betas, SWEEPS, T=T, seed=1000 + k,
                                mode=oracle.MODE_REFFP)["best_cost"] for k in range(RUNS)], float)
    ex = np.array([oracle.solve(F, D, betas, SWEEPS, T=T, seed=2000 + k,
                                mode=oracle.MODE_EXACT)["best_cost"] for k in range(RUNS)], float)
    for mine in (fp, ex):
        se = np.sqrt(ref.var(ddof=1) / RUNS + mine.var(ddof=1) / RUNS)
        gap = abs(ref.mean() - mine.mean())
        assert gap <= 5 * se + 0.002 * ref.mean(), (ref.mean(), mine.mean(), se)
    # and both are far from a random permutation's cost: the search is doing work
    rand = np.mean([qpt.instances.qap_cost(F, D, np.random.default_rng(k).permutation(N))
                    for k in range(50)])
    assert ref.mean() < 0.9 * rand and ex.mean() < 0.9 * rand
