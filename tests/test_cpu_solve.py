"""qap_pt_solve_cpu (the product's host solve, CPU_PIM.cpp:478-685) against the oracle.

The host solve runs the device chain on the host's cores; here it is compared bit for
bit with oracle/qap_oracle.c's exact mode (best cost and permutation, accepted moves,
exchanges, kicks, attempts, sweeps done) on the instance families the device parity
cases use, at several thread counts.  No GPU is needed: the call makes no device call.
"""
import numpy as np
import pytest


def _wide64(seed, n, symmetric=True):
    """Flow entries up to 2^27: force bounds beyond int32 (the int64 force matrices)."""
    g = np.random.default_rng(seed)
    F = g.integers(0, 1 << 27, (n, n), dtype=np.int64)
    D = g.integers(0, 64, (n, n), dtype=np.int64)
    if symmetric:
        F = np.triu(F, 1) + np.triu(F, 1).T
        D = np.triu(D, 1) + np.triu(D, 1).T
    np.fill_diagonal(F, 0)
    np.fill_diagonal(D, 0)
    return F.astype(np.int32), D.astype(np.int32)


def _sparse_asym(seed, n, density):
    g = np.random.default_rng(seed)
    F = np.where(g.random((n, n)) < density, g.integers(1, 40, (n, n)), 0).astype(np.int32)
    D = g.integers(0, 30, (n, n)).astype(np.int32)
    return F, D


CASES = {
    # name: (instance builder, R, sweeps, kwargs)
    "sym_n16_R4": (lambda q: q.instances.random_instance(11, 16), 4, 40, {}),
    "c1_n64_R16": (lambda q: q.instances.flat_c1(), 16, 30, {}),
    "asym_diag_n24_T2_spe2": (lambda q: q.instances.random_instance(
        12, 24, symmetric=False, zero_diag=False), 6, 20, {"T": 2, "spe": 2}),
    "asym_n70_R5": (lambda q: q.instances.random_instance(13, 70, 0, 50, symmetric=False), 5, 6,
                    {}),
    "sparse_n150_R4": (lambda q: q.instances.mixed_instance(6, 150, False), 4, 3, {}),
    "sparse_asym_n60": (lambda q: _sparse_asym(40, 60, 0.04), 4, 10, {}),
    "grey_n256_R3": (lambda q: q.instances.grey_c2(), 3, 2, {}),
    "per_replica_init_n33_T3": (lambda q: q.instances.random_instance(15, 33), 5, 10,
                                {"T": 3, "per_replica_init": True}),
    "kick_often_n20": (lambda q: q.instances.random_instance(16, 20), 4, 60, {"kick_after": 2}),
    "ladders_T12_kick": (lambda q: q.instances.random_instance(32, 24), 4, 30,
                         {"T": 12, "kick_after": 3}),
    "n1": (lambda q: q.instances.random_instance(27, 1), 3, 5, {}),
    "n2": (lambda q: q.instances.random_instance(28, 2), 2, 20, {}),
    "wide64_sym_n40": (lambda q: _wide64(80, 40), 4, 10, {"beta_scale": "wide"}),
    "wide64_asym_n37_T2": (lambda q: _wide64(81, 37, symmetric=False), 3, 8,
                           {"T": 2, "beta_scale": "wide", "kick_after": 2}),
}


def _betas(qpt, R, n, scale=None):
    b = qpt.instances.geometric_ladder(R, 1.0, 100.0) / max(1.0, n / 16.0)
    if scale == "wide":  # deltas near 4 * 63 * 2^27 * n
        b = b * 20.0 / (4 * 63 * (1 << 27) * n) * max(1.0, n / 16.0)
    return b


def _instance(qpt, F, D, betas, sweeps, T=1, spe=1, seed=7, kick_after=100, goal=None,
              per_replica_init=False, init_perm=None):
    return qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=sweeps, T=T, sweeps_per_exchange=spe,
                        seed=seed, kick_after=kick_after, goal=goal,
                        per_replica_init=per_replica_init, init_perm=init_perm)


def _assert_matches_oracle(res, ref):
    assert res["best_cost"] == ref["best_cost"]
    np.testing.assert_array_equal(res["best_perm"], ref["best_perm"])
    assert res["accepted"] == ref["accepted"]
    assert res["exchanges"] == ref["exchanges"]
    assert res["kicks"] == ref["kicks"]
    assert res["attempts"] == ref["attempts"]


@pytest.mark.parametrize("name", list(CASES))
def test_cpu_solve_matches_oracle(qpt, oracle, name):
    build, R, sweeps, kw = CASES[name]
    kw = dict(kw)
    F, D = build(qpt)
    n = F.shape[0]
    betas = _betas(qpt, R, n, kw.pop("beta_scale", None))
    res = qpt.solve_cpu(_instance(qpt, F, D, betas, sweeps, seed=7, **kw), threads=3)
    ref = oracle.solve(F, D, betas, sweeps, seed=7, **kw)
    _assert_matches_oracle(res, ref)
    assert res["sweeps_done"] == sweeps


def test_cpu_solve_thread_count_invariant(qpt):
    """Replicas split over 1, 2, 5 or 16 threads give the same result, counters included."""
    F, D = qpt.instances.random_instance(29, 40)
    betas = _betas(qpt, 8, 40)
    inst = _instance(qpt, F, D, betas, 12, T=2)
    outs = [qpt.solve_cpu(inst, threads=t) for t in (1, 2, 5, 16)]
    for o in outs[1:]:
        for k in ("best_cost", "accepted", "exchanges", "kicks", "attempts", "update_cells",
                  "sweeps_done"):
            assert o[k] == outs[0][k], k
        np.testing.assert_array_equal(o["best_perm"], outs[0]["best_perm"])


def test_cpu_solve_goal_stops_like_oracle(qpt, oracle):
    F, D = qpt.instances.random_instance(21, 30)
    betas = qpt.instances.geometric_ladder(6, 1.0, 50.0)
    free = oracle.solve(F, D, betas, 60, seed=3)
    goal = int(free["best_cost"] * 1.02)
    ref = oracle.solve(F, D, betas, 60, seed=3, goal=goal)
    res = qpt.solve_cpu(_instance(qpt, F, D, betas, 60, seed=3, goal=goal), threads=4)
    assert ref["megas_done"] < 60
    assert res["sweeps_done"] == ref["megas_done"]
    _assert_matches_oracle(res, ref)


def test_cpu_solve_init_perm(qpt, oracle):
    F, D = qpt.instances.random_instance(22, 30)
    betas = _betas(qpt, 4, 30)
    ip = np.random.default_rng(5).permutation(30).astype(np.int32)
    res = qpt.solve_cpu(_instance(qpt, F, D, betas, 15, seed=7, init_perm=ip), threads=2)
    ref = oracle.solve(F, D, betas, 15, seed=7, init_perm=ip)
    _assert_matches_oracle(res, ref)


def test_cpu_solve_update_cells_count(qpt):
    """update_cells is n times the nonzero flow-difference columns of every accepted
    move.  With F[x][y] = x + y off the diagonal (0 on it), a swap of a and b changes
    every column: b - a outside {a, b}, a + b at column a and -(a + b) at column b, so
    cells = n^2 x accepted exactly.  With one nonzero flow pair, a move changes at most
    the two columns 0 and 1."""
    n = 12
    x = np.arange(n)
    F = (x[:, None] + x[None, :]).astype(np.int32)
    np.fill_diagonal(F, 0)
    D = np.triu(np.random.default_rng(3).integers(1, 9, (n, n)), 1)
    D = (D + D.T).astype(np.int32)
    betas = _betas(qpt, 3, n)
    res = qpt.solve_cpu(_instance(qpt, F, D, betas, 8), threads=2)
    assert res["accepted"] > 0
    assert res["update_cells"] == n * n * res["accepted"]
    P = np.zeros((n, n), np.int32)
    P[0, 1] = P[1, 0] = 5
    res = qpt.solve_cpu(_instance(qpt, P, D, np.full(2, 1e-3), 8), threads=1)
    assert res["update_cells"] % n == 0
    assert 0 < res["update_cells"] <= 2 * n * res["accepted"]


def test_cpu_solve_errors(qpt):
    F, D = qpt.instances.random_instance(23, 10)
    betas = _betas(qpt, 3, 10)
    with pytest.raises(qpt.QapError) as e:
        qpt.solve_cpu(qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=5, one_shot=True))
    assert "device only" in str(e.value)
    with pytest.raises(qpt.QapError):
        qpt.solve_cpu(qpt.Instance(F=F, D=D, beta_ln=betas[::-1].copy(), sweeps=5))
    with pytest.raises(qpt.QapError):
        qpt.solve_cpu(qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=5, sweeps_per_exchange=2))
    big = np.full((10, 10), 2_000_000_000, np.int32)
    np.fill_diagonal(big, 0)
    with pytest.raises(qpt.QapError) as e:
        qpt.solve_cpu(qpt.Instance(F=big, D=big, beta_ln=betas, sweeps=5))
    assert e.value.code == -4  # QAP_PT_E_OVERFLOW
