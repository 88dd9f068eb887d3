"""GPU: the automatic ladder scan and the command-line front end against the oracle.

The scan's 32 x 5 trials are single instances of the batched pool; each trial's final
replica costs must equal the oracle's for the same seed and ladder, so the averages (summed
in the same order) and the chosen ladder are exact.  The command line must print the
oracle's best cost for the ladder its -b/-m flags describe.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG_DIR, "lib", "qap_pt_cli")


def _trial_seed(seed, i, t):
    return (seed ^ (0x9E3779B97F4A7C15 * (1 + i * 5 + t))) & 0xFFFFFFFFFFFFFFFF


def test_auto_beta_matches_oracle_trials(qpt, oracle, gpu):
    F, D = qpt.instances.random_instance(31, 14)
    R, T, spe, seed = 4, 2, 2, 77
    print(f"auto_beta n=14 R={R} T={T} spe={spe}", flush=True)
    b0, m, avg = qpt.auto_beta(F, D, R, T=T, sweeps_per_exchange=spe, seed=seed)
    ref = np.zeros(32)
    for i in range(32):
        lad = np.full(R, 2.0 ** (-30 + i) / 1.44269504089)
        for t in range(5):
            r = oracle.solve(F, D, lad, 10 * spe, T=T, spe=spe, seed=_trial_seed(seed, i, t))
            for e in r["final_E"]:
                ref[i] += float(e) / (5.0 * T * R)
    np.testing.assert_array_equal(avg, ref)
    assert (b0, m) == qpt.auto_select(ref, R)
    # the scan itself: random costs at the hottest beta, far lower at the coldest
    assert avg[-1] < avg[0]


def _write_dat(path, F, D):
    n = F.shape[0]
    with open(path, "w") as fh:
        fh.write(f"{n}\n\n")
        for M in (F, D):
            for row in M:
                fh.write(" ".join(str(int(x)) for x in row) + "\n")
            fh.write("\n")


def test_cli_best_cost_matches_oracle(qpt, oracle, gpu, tmp_path):
    F, D = qpt.instances.random_instance(32, 20)
    path = tmp_path / "r20.dat"
    _write_dat(path, F, D)
    out = tmp_path / "out"
    R, T, w, i, seed = 6, 3, 150, 2, 4242
    cmd = [CLI, "-b", "0.05", "-m", "1.4", "-r", str(R), "-t", str(T), "-w", str(w), "-i",
           str(i), "--seed", str(seed), "-G", "1000000000000", "-o", str(out), str(path)]
    print("cli:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    best = int(re.search(r"Single best energy found: (-?\d+)", r.stdout).group(1))
    betas = qpt.ladder_from(0.05, 1.4, R)
    ref = oracle.solve(F, D, betas, w * i, T=T, spe=i, seed=seed)
    assert best == ref["best_cost"]
    # -G prints the best permutation; its cost must be the reported one
    lines = r.stdout.splitlines()
    k = next(x for x, line in enumerate(lines) if line.startswith("Halting"))
    perm = np.array([int(v) for v in lines[k - 2].split()])
    assert qpt.instances.qap_cost(F, D, perm) == best
    log = (out / "log.txt").read_text()
    assert "Number of Replicas 6" in log and f"Single best energy found: {best}" in log
    # one sweep<t>.txt per ladder (CPU_PIM.cpp:320,660-664): completed exchange intervals,
    # milliseconds and the ladder's best, improving line by line; the last line is the
    # best over that ladder's replicas at the end of the run
    for t in range(T):
        rows = [[int(v) for v in line.split()]
                for line in (out / f"sweep{t}.txt").read_text().splitlines()]
        assert rows, t
        assert all(a[2] > b[2] and a[0] < b[0] for a, b in zip(rows, rows[1:])), rows
        assert all(0 < r[0] <= w for r in rows)
        assert rows[-1][2] == int(np.min(ref["final_bestE"][t * R:(t + 1) * R]))
    assert not (out / f"sweep{T}.txt").exists()


def test_cli_automatic_mode_runs(qpt, gpu, tmp_path):
    F, D = qpt.instances.random_instance(33, 12)
    path = tmp_path / "r12.dat"
    _write_dat(path, F, D)
    cmd = [CLI, "-r", "4", "-t", "2", "-w", "50", "-o", str(tmp_path / "o"), str(path)]
    print("cli:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Average Energy") == 32
    assert "Selecting minimum beta" in r.stdout and "Selecting maximum beta" in r.stdout
    best = int(re.search(r"Single best energy found: (-?\d+)", r.stdout).group(1))
    assert best >= 0


def test_cli_one_shot_matches_oracle(qpt, oracle, gpu, tmp_path):
    """--one-shot runs GPU-CUPIM's dynamics: the printed best is the oracle's one-shot run
    for the same ladder, and the automatic scan's trials use the same dynamics."""
    F, D = qpt.instances.random_instance(34, 18)
    path = tmp_path / "r18.dat"
    _write_dat(path, F, D)
    R, T, w, i, seed = 5, 2, 60, 3, 99
    cmd = [CLI, "--one-shot", "-b", "0.05", "-m", "1.4", "-r", str(R), "-t", str(T), "-w",
           str(w), "-i", str(i), "--seed", str(seed), "-o", str(tmp_path / "o"), str(path)]
    print("cli:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    best = int(re.search(r"Single best energy found: (-?\d+)", r.stdout).group(1))
    ref = oracle.solve(F, D, qpt.ladder_from(0.05, 1.4, R), w * i, T=T, spe=i, seed=seed,
                       one_shot=True)
    assert best == ref["best_cost"]
    # automatic mode with one-shot trials: the scan's averages are the oracle's one-shot runs
    b0, m, avg = qpt.auto_beta(F, D, 3, T=1, sweeps_per_exchange=1, seed=5, one_shot=True)
    lad = np.full(3, 2.0 ** (-30 + 20) / 1.44269504089)
    want = 0.0
    for t in range(5):
        rr = oracle.solve(F, D, lad, 10, T=1, spe=1, seed=_trial_seed(5, 20, t), one_shot=True)
        for e in rr["final_E"]:
            want += float(e) / (5.0 * 1 * 3)
    assert avg[20] == want
