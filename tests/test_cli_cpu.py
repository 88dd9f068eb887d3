"""The command line with --cpu (qap_pt_solve_cpu, no device call): CPU_PIM's flags and log
lines on the host's cores, with the oracle's best cost for the ladder -b/-m describe."""
import os
import re
import subprocess

import numpy as np

from conftest import PKG_DIR

CLI = os.path.join(PKG_DIR, "lib", "qap_pt_cli")


def _write_dat(path, F, D):
    n = F.shape[0]
    with open(path, "w") as fh:
        fh.write(f"{n}\n\n")
        for M in (F, D):
            for row in M:
                fh.write(" ".join(str(int(x)) for x in row) + "\n")
            fh.write("\n")


def test_cli_cpu_matches_oracle(qpt, oracle, tmp_path):
    F, D = qpt.instances.random_instance(32, 20)
    path = tmp_path / "r20.dat"
    _write_dat(path, F, D)
    out = tmp_path / "out"
    R, T, w, i, seed = 6, 3, 150, 2, 4242
    cmd = [CLI, "--cpu", "-b", "0.05", "-m", "1.4", "-r", str(R), "-t", str(T), "-w", str(w),
           "-i", str(i), "--seed", str(seed), "-G", "1000000000000", "-o", str(out), str(path)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, QAP_PT_CPU_THREADS="4"))
    assert r.returncode == 0, r.stderr
    best = int(re.search(r"Single best energy found: (-?\d+)", r.stdout).group(1))
    ref = oracle.solve(F, D, qpt.ladder_from(0.05, 1.4, R), w * i, T=T, spe=i, seed=seed)
    assert best == ref["best_cost"]
    lines = r.stdout.splitlines()
    k = next(x for x, line in enumerate(lines) if line.startswith("Halting"))
    perm = np.array([int(v) for v in lines[k - 2].split()])
    assert qpt.instances.qap_cost(F, D, perm) == best
    log = (out / "log.txt").read_text()
    assert f"Best Energy Found: {best} Time (s):" in log
    assert f"Single best energy found: {best}" in log and "Elapsed time:" in log


def test_cli_cpu_needs_an_explicit_ladder(qpt, tmp_path):
    F, D = qpt.instances.random_instance(33, 8)
    path = tmp_path / "r8.dat"
    _write_dat(path, F, D)
    for extra in ([], ["-b", "0.1", "--one-shot"]):
        r = subprocess.run([CLI, "--cpu", *extra, "-o", str(tmp_path / "o"), str(path)],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "--cpu" in r.stderr, (extra, r.stderr)
