"""qap_pt_solve_cpu against qap_pt_solve_hip on the device: the host and device solves are
the same chain, so every result field agrees exactly, update_cells included (the count the
sweep kernels accumulate per wave), on the placements the device picks by size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("best_cost", "accepted", "exchanges", "kicks", "attempts", "update_cells",
        "sweeps_done")


def _wide64(seed, n):
    g = np.random.default_rng(seed)
    F = np.triu(g.integers(0, 1 << 27, (n, n), dtype=np.int64), 1)
    D = np.triu(g.integers(0, 64, (n, n), dtype=np.int64), 1)
    return (F + F.T).astype(np.int32), (D + D.T).astype(np.int32)


CASES = {
    # name: (instance builder, R, sweeps, kwargs) -- device placement in the comment
    "c1_n64_R16": (lambda q: q.instances.flat_c1(), 16, 40, {}),                 # lds32
    "grey_n256_R8": (lambda q: q.instances.grey_c2(), 8, 4, {"ladder": (10.0, 1000.0)}),  # lds16
    "asym_n70_T2": (lambda q: q.instances.random_instance(13, 70, 0, 50, symmetric=False), 4, 6,
                    {"T": 2}),                                                   # lds32 general
    "sparse_n150": (lambda q: q.instances.mixed_instance(6, 150, False), 4, 4, {}),
    "c5_n512_R4": (lambda q: q.instances.large_c5(512), 4, 1, {}),              # lazy
    "c5_n1024_R2": (lambda q: q.instances.large_c5(1024), 2, 1, {}),            # lazy + helper
    "wide64_n100": (lambda q: _wide64(82, 100), 3, 4, {"wide": True}),          # hbm64
}


@pytest.mark.parametrize("name", list(CASES))
def test_cpu_solve_equals_device_solve(qpt, gpu, name):
    build, R, sweeps, kw = CASES[name]
    kw = dict(kw)
    F, D = build(qpt)
    n = F.shape[0]
    ladder = kw.pop("ladder", None)
    if kw.pop("wide", False):
        betas = (qpt.instances.geometric_ladder(R, 1.0, 100.0) * 20.0
                 / (4 * 63 * (1 << 27) * n))
    elif ladder is not None:
        betas = qpt.instances.geometric_ladder(R, *ladder)
    else:
        betas = qpt.instances.geometric_ladder(R, 1.0, 100.0) / max(1.0, n / 16.0)
    print(f"case {name}: n={n} R={R} sweeps={sweeps} kw={kw}", flush=True)
    inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=sweeps, seed=11, **kw)
    dev = qpt.solve(inst)
    host = qpt.solve_cpu(inst, threads=8)
    for k in KEYS:
        assert host[k] == dev[k], (k, host[k], dev[k])
    np.testing.assert_array_equal(host["best_perm"], dev["best_perm"])
