"""The allocation-free solve (qap_pt_solve_hip_noalloc) against the allocating one.

Both calls go through the C-ABI.  Two calls in a row, the second on a smaller instance,
must each equal qap_pt_solve_hip's result bit for bit (the region keeps the first call's
contents, so the second shows that nothing stale is read); the results also equal the
oracle's.  A request larger than the fixed region is refused before any device work.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    for key in ("best_cost", "attempts", "accepted", "exchanges", "kicks", "update_cells",
                "flush_cells", "sweeps_done"):
        assert a[key] == b[key], (key, a[key], b[key])
    np.testing.assert_array_equal(a["best_perm"], b["best_perm"])


def test_noalloc_equals_allocating_solve_twice(qpt, oracle, gpu):
    cases = [
        # first a larger instance (dense flow, lazy-free int32 LDS placement, 2 ladders)
        (qpt.instances.mixed_instance(201, 120, True), 16, 2, 40, 11),
        # then a smaller one (sparse flow): every piece it reads is rewritten first
        (qpt.instances.mixed_instance(202, 45, False), 8, 1, 30, 12),
    ]
    for (F, D), R, T, sweeps, seed in cases:
        betas = qpt.instances.geometric_ladder(R, 1.0, 100.0) / 4.0
        inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=sweeps, T=T, seed=seed)
        print(f"noalloc case n={inst.n} R={R} T={T} sweeps={sweeps}", flush=True)
        got = qpt.solve_noalloc(inst)
        want = qpt.solve(inst)
        _same(got, want)
        ref = oracle.solve(F, D, betas, sweeps, T=T, seed=seed)
        assert got["best_cost"] == ref["best_cost"]
        np.testing.assert_array_equal(got["best_perm"], ref["best_perm"])
        assert got["best_cost"] == qpt.instances.qap_cost(F, D, got["best_perm"])


def test_noalloc_grey_pattern_n256(qpt, gpu):
    """The grey-pattern n = 256 instance with 64 replicas (about 18 MB of device state)
    in the uint16 LDS placement, twice, against the allocating solve."""
    F, D = qpt.instances.grey_c2()
    betas = qpt.instances.geometric_ladder(64, 10.0, 1000.0)
    inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=6, seed=2024)
    want = qpt.solve(inst)
    for _ in range(2):
        _same(qpt.solve_noalloc(inst), want)


def test_noalloc_refuses_what_the_region_cannot_hold(qpt, gpu):
    """n = 1024 with 64 replicas needs about 270 MB: QAP_PT_E_NOMEM, nothing launched, and
    the next call still works."""
    F, D = qpt.instances.large_c5(1024)
    big = qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(64, 300.0, 30000.0),
                       sweeps=1, seed=1)
    with pytest.raises(qpt.QapError) as ei:
        qpt.solve_noalloc(big)
    assert ei.value.code == qpt.QAP_PT_E_NOMEM
    assert "bytes" in str(ei.value)
    F, D = qpt.instances.random_instance(7, 20)
    small = qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(4, 1.0, 10.0),
                         sweeps=10, seed=3)
    _same(qpt.solve_noalloc(small), qpt.solve(small))
