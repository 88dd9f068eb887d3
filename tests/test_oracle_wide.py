"""The oracle's wide path: instances whose force bound exceeds int32 run with int64 force
matrices and deltas (CUPIM's ReplicaPool<F, int64_t>, GPU-CUPIM/perl-gpu/src/lib/qap/
qap.cu:369-411).  Pinned without restating its arithmetic: every tracked cost equals the
cost recomputed from scratch in Python integers, the bests are the costs of the best
permutations, small instances reach the brute-force optimum, and the 64-bit weight equals
the 32-bit one wherever both apply (so the int32 and wide paths share one definition)."""
import itertools

import numpy as np
import pytest

from conftest import load_qpt


def _cost(F, D, p):
    n = len(p)
    return sum(int(D[i, j]) * int(F[p[i], p[j]]) for i in range(n) for j in range(n))


def _wide(seed, n, symmetric, big=1 << 27):
    g = np.random.default_rng(seed)
    F = g.integers(0, big, (n, n), dtype=np.int64)
    D = g.integers(0, 64, (n, n), dtype=np.int64)
    if symmetric:
        F = np.triu(F, 1) + np.triu(F, 1).T
        D = np.triu(D, 1) + np.triu(D, 1).T
    return F.astype(np.int32), D.astype(np.int32)


def _is_wide(oracle, F, D):
    fb, db = oracle.bounds(F, D)
    return db >= (1 << 30)


@pytest.mark.parametrize("sym,n,T,spe", [(True, 14, 1, 1), (False, 11, 2, 2), (True, 30, 1, 3)])
def test_wide_costs_equal_recomputed(oracle, sym, n, T, spe):
    F, D = _wide(40 + n, n, sym)
    assert _is_wide(oracle, F, D)
    dmax = 4 * 64 * (1 << 27) * n
    betas = load_qpt().instances.geometric_ladder(5, 1.0, 100.0) * (20.0 / dmax)
    for mode in (oracle.MODE_EXACT, oracle.MODE_REFFP):
        r = oracle.solve(F, D, betas, 24, T=T, spe=spe, seed=5, mode=mode)
        assert r["accepted"] > 0
        for k in range(len(r["final_E"])):
            assert r["final_E"][k] == _cost(F, D, r["final_perm"][k])
            if r["final_bestE"][k] != np.iinfo(np.int64).max:
                assert r["final_bestE"][k] == _cost(F, D, r["best_perms"][k])
        assert r["best_cost"] == _cost(F, D, r["best_perm"])
        assert r["best_cost"] > (1 << 31)  # costs beyond int32


@pytest.mark.parametrize("n,sym", [(6, True), (7, False)])
def test_wide_reaches_brute_force_optimum(oracle, n, sym):
    F, D = _wide(70 + n, n, sym)
    assert _is_wide(oracle, F, D)
    best = min(_cost(F, D, p) for p in itertools.permutations(range(n)))
    dmax = 4 * 64 * (1 << 27) * n
    betas = load_qpt().instances.geometric_ladder(6, 0.05, 5.0) * (20.0 / dmax)
    r = oracle.solve(F, D, betas, 300, seed=3)
    assert r["best_cost"] == best


def test_weight64_equals_weight_in_int32_range(oracle):
    g = np.random.default_rng(9)
    for _ in range(2000):
        b = float(np.float32(10 ** g.uniform(-9, 1)))
        d = -int(g.integers(0, 1 << 31))
        assert oracle.weight64(b, d) == oracle.weight(b, d)
    # beyond int32 the weight follows the same float steps from y = beta * (float)d on
    assert oracle.weight64(1e-10, -(1 << 40)) == oracle.weight(1e-10 * float(1 << 20), -(1 << 20))
    assert oracle.weight64(1.0, -(1 << 40)) == 0


def test_entries_beyond_int64_are_refused(oracle):
    F, D = _wide(3, 8, True, big=(1 << 31) - 1)
    D = (D.astype(np.int64) * (1 << 25)).clip(0, (1 << 31) - 1).astype(np.int32)
    with pytest.raises(ValueError):
        oracle.solve(F, D, np.array([1e-20, 2e-20]), 2)
