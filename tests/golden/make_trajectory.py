"""Writes tests/golden/trajectory_n24.json: a seeded PT-SA trajectory of the CPU oracle
(exact mode) -- energies of every replica after every exchange interval, the final
permutations and the result.  Regenerate only when the trajectory definition changes
on purpose; tests/test_oracle.py and tests/test_gpu_parity.py pin against it."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))

import numpy as np  # noqa: E402
import pyoracle  # noqa: E402
from conftest import load_qpt  # noqa: E402

qpt = load_qpt()
F, D = qpt.instances.random_instance(2024, 24, 0, 30, symmetric=False, zero_diag=False)
betas = qpt.instances.geometric_ladder(6, 2.0, 400.0)
cfg = dict(sweeps=40, T=2, seed=0xABCDEF0123)
r = pyoracle.solve(F, D, betas, cfg["sweeps"], T=cfg["T"], seed=cfg["seed"], traj=True)
out = {
    "F": F.tolist(), "D": D.tolist(), "beta_ln": betas.tolist(), **cfg,
    "traj_E": r["traj_E"].tolist(), "final_perm": r["final_perm"].tolist(),
    "final_pos": r["final_pos"].tolist(), "best_cost": r["best_cost"],
    "best_perm": r["best_perm"].tolist(), "accepted": r["accepted"],
    "exchanges": r["exchanges"], "kicks": r["kicks"],
}
with open(os.path.join(HERE, "trajectory_n24.json"), "w") as fh:
    json.dump(out, fh, separators=(",", ":"))
print("best", r["best_cost"], "accepted", r["accepted"])
