"""Pins of the CPU oracle (oracle/qap_oracle.c) -- CPU only.

Each check is chosen so that a plausible mistake in the restatement (a dropped term,
a wrong sign or index, a transposed operand, an off-by-one in the roulette) fails:
  * Philox4x32-10: Random123 known-answer vectors and 64 vectors from rocRAND's
    philox4x32_10 engine (tests/golden/philox_rocrand.json).
  * Force matrix: the gradient identity -I[i][a] = dE/dX[i][a] of the quadratic form
    E(X) = sum D[i][j] F[a][b] X[i][a] X[j][b] at a permutation matrix, and
    2 E(p) = -sum_i I[i][p(i)].
  * O(1) delta: equals E(p) - E(p') recomputed from scratch for every (i, b).
  * Rank-2 update: equals a rebuild of I for the swapped permutation.
  * Heat-bath weights/selection: exact Gibbs probabilities 2^(beta dE) (within the
    2^-20 resolution), the beta -> 0 and beta -> infinity limits, and the inverse-CDF
    selection over a uniform grid of random words.
  * PT driver: hand-built exchange/kick vectors restating CPU_PIM.cpp:323-344.
  * Whole solve: brute-force optimum for n <= 8; the committed golden trajectory.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# Random123 kat_vectors for philox4x32_10 (counter, key -> output).
R123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff, 0xffffffff], [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_random123_kat(oracle):
    for ctr, key, out in R123_KAT:
        assert oracle.philox(ctr, key) == out


def test_philox_matches_rocrand_engine(oracle):
    with open(os.path.join(GOLDEN, "philox_rocrand.json")) as fh:
        vecs = json.load(fh)
    assert len(vecs) == 64
    for v in vecs:
        assert oracle.philox(v["ctr"], v["key"]) == v["out"]


def _perm_matrix(p):
    n = len(p)
    X = np.zeros((n, n), dtype=np.int64)
    X[np.arange(n), p] = 1
    return X


def _quad_energy(F, D, X):
    # E(X) = sum_{i,j,a,b} D[i][j] F[a][b] X[i][a] X[j][b] = sum(D * (X F X^T))
    return int((D.astype(np.int64) * (X @ F.astype(np.int64) @ X.T)).sum())


@pytest.mark.parametrize("symmetric,zero_diag", [(True, True), (False, False), (False, True)])
def test_force_matrix_is_gradient_of_quadratic_form(qpt, oracle, symmetric, zero_diag):
    rng = np.random.default_rng(3)
    for trial in range(6):
        n = int(rng.integers(2, 10))
        F, D = qpt.instances.random_instance(100 + trial, n, -5, 9, symmetric, zero_diag)
        p = rng.permutation(n).astype(np.int32)
        I = oracle.force_matrix(F, D, p)
        X = _perm_matrix(p)
        E0 = _quad_energy(F, D, X)
        for i in range(n):
            for a in range(n):
                Xa = X.copy()
                Xa[i, a] += 1
                grad = _quad_energy(F, D, Xa) - E0 - int(D[i, i]) * int(F[a, a])
                assert I[i, a] == -grad
        assert 2 * oracle.energy(F, D, p) == -int(I[np.arange(n), p].sum())
        assert oracle.energy(F, D, p) == qpt.instances.qap_cost(F, D, p) == E0


@pytest.mark.parametrize("symmetric,zero_diag", [(True, True), (False, False)])
def test_delta_and_rank2_update_match_recompute(qpt, oracle, symmetric, zero_diag):
    rng = np.random.default_rng(4)
    for trial in range(5):
        n = int(rng.integers(2, 12))
        F, D = qpt.instances.random_instance(200 + trial, n, -7, 9, symmetric, zero_diag)
        p = rng.permutation(n).astype(np.int32)
        I = oracle.force_matrix(F, D, p)
        E = oracle.energy(F, D, p)
        pinv = np.argsort(p)
        for i in range(n):
            for b in range(n):
                j = int(pinv[b])
                a = int(p[i])
                q = p.copy()
                q[i], q[j] = b, a
                assert oracle.delta(F, D, p, I, i, b) == E - oracle.energy(F, D, q)
                if b != a:
                    I2 = oracle.force_update(F, D, I, i, j, a, b)
                    np.testing.assert_array_equal(I2, oracle.force_matrix(F, D, q))


def test_weight_exact_points_and_resolution(oracle):
    assert oracle.weight(1.0, 0) == 1 << 20
    assert oracle.weight(0.5, 0) == 1 << 20
    for k in range(0, 21):
        assert oracle.weight(1.0, -k) == 1 << (20 - k)   # 2^(20-k) exactly
    assert oracle.weight(1.0, -21) == 0
    assert oracle.weight(1.0, -1000) == 0
    rng = np.random.default_rng(5)
    for _ in range(4000):
        beta = float(np.float32(10 ** rng.uniform(-4, 1)))
        d = -int(rng.integers(0, 2000))
        y = float(np.float32(np.float32(beta) * np.float32(d)))
        exact = math.floor(2.0 ** (20 + y)) if y > -32 else 0
        w = oracle.weight(beta, d)
        assert abs(w - exact) <= max(1, exact * 2e-7), (beta, d, w, exact)


def test_weight_monotone_in_delta(oracle):
    for beta in (0.013, 0.5, 1.4427, 7.0):
        prev = oracle.weight(beta, 0)
        for d in range(-1, -400, -1):
            w = oracle.weight(beta, d)
            assert w <= prev
            prev = w


def _state(qpt, oracle, seed, n):
    F, D = qpt.instances.random_instance(seed, n, 0, 9)
    p = np.random.default_rng(seed).permutation(n).astype(np.int32)
    return F, D, p, oracle.force_matrix(F, D, p), oracle.energy(F, D, p)


def test_selection_probabilities_are_gibbs(qpt, oracle):
    """Frequencies of the chosen partner over a uniform grid of random words equal
    w_b / S, and w_b / S is the heat-bath probability 2^(beta dE_b) / Z up to the 2^-20
    weight resolution (candidates below the reference cutoff excluded)."""
    F, D, p, I, E = _state(qpt, oracle, 7, 9)
    n = 9
    for beta_ln in (0.002, 0.02, 0.2):
        b2 = oracle.beta2(beta_ln)
        for i in range(n):
            _, dE, w, _, _, _ = oracle.attempt_exact(F, D, p, I, E, i, np.float32(b2), 20.0 / b2, 0)
            S = int(w.sum())
            dEf = dE.astype(np.float64)
            cut = int(max(0, dE.max()) - 20.0 / b2)
            keep = dE > cut
            z = np.where(keep, 2.0 ** (b2 * (dEf - max(0, dE.max()))), 0.0)
            gibbs = z / z.sum()
            np.testing.assert_allclose(w / S, gibbs, atol=4 * n / 2 ** 20)
            counts = np.zeros(n)
            grid = 4096
            for k in range(grid):
                u = (k * (1 << 32)) // grid
                sel = oracle.attempt_exact(F, D, p, I, E, i, np.float32(b2), 20.0 / b2, u)[0]
                counts[sel] += 1
            np.testing.assert_allclose(counts / grid, w / S, atol=1.5 / grid)


def test_selection_limits(qpt, oracle):
    F, D, p, I, E = _state(qpt, oracle, 8, 10)
    # beta -> 0: every candidate (including "stay") almost equally likely
    b2 = oracle.beta2(1e-9)
    _, dE, w, _, _, _ = oracle.attempt_exact(F, D, p, I, E, 3, np.float32(b2), 20.0 / b2, 0)
    assert w.min() >= (1 << 20) - 2 and w.max() == 1 << 20
    # beta -> infinity: only the candidates at the maximum delta keep a weight, and
    # the selection is the first of them for the smallest random word
    b2 = oracle.beta2(1e3)
    for i in range(10):
        for u in (0, 1 << 31, (1 << 32) - 1):
            sel, dE, w, _, _, _ = oracle.attempt_exact(F, D, p, I, E, i, np.float32(b2),
                                                       20.0 / b2, u)
            top = max(0, int(dE.max()))
            if top > 0:
                np.testing.assert_array_equal(w > 0, dE == top)
                assert dE[sel] == top
                if u == 0:
                    assert sel == int(np.argmax(dE))


def test_all_weights_zero_means_stay(qpt, oracle):
    """maxd = 0 and 20/beta < 1: the reference cutoff (int)(0 - 20/beta) = 0 rejects
    every candidate; the oracle stays (the reference would pick b = 0)."""
    F, D = qpt.instances.random_instance(9, 6, 0, 9)
    # a local optimum: run a greedy descent to dE <= 0 for location 2
    p = np.arange(6, dtype=np.int32)
    for _ in range(50):
        I = oracle.force_matrix(F, D, p)
        E = oracle.energy(F, D, p)
        b2 = oracle.beta2(1e3)
        sel, dE, w, p2, _, _ = oracle.attempt_exact(F, D, p, I, E, 2, np.float32(b2), 20.0 / b2, 0)
        if dE.max() <= 0:
            break
        p = p2
    I = oracle.force_matrix(F, D, p)
    E = oracle.energy(F, D, p)
    assert oracle.attempt_exact(F, D, p, I, E, 2, np.float32(b2), 20.0 / b2, 0)[1].max() == 0
    sel, dE, w, p3, _, E3 = oracle.attempt_exact(F, D, p, I, E, 2, np.float32(b2), 20.0 / b2,
                                                 12345)
    assert w.sum() == 0 and sel == p[2] and E3 == E
    np.testing.assert_array_equal(p3, p)


def test_exchange_step_hand_vectors(oracle):
    """CPU_PIM.cpp:323-344: hottest position first; swap when the hotter rung has the
    lower energy; bubbling carries a replica several positions; timeSinceSwap resets
    only when the *replica index* R-1 moves up; kick swaps positions R-1 and 0."""
    R = 4
    order = np.array([0, 1, 2, 3])
    pos = np.array([0, 1, 2, 3])
    big = np.iinfo(np.int64).max
    # replica 0 (hottest) has the lowest energy: it bubbles to the coldest position;
    # its last swap is with replica index R-1, which resets timeSinceSwap
    s = oracle.exchange_step([10, 50, 40, 30], [big] * R, order, pos, tss=5, ladder_best=big)
    assert list(s["order"]) == [1, 2, 3, 0]
    assert list(s["pos"]) == [3, 0, 1, 2]
    assert s["exchanges"] == 3 and s["tss"] == 0 and not s["kicked"]
    # a swap that does not involve replica R-1 leaves the counter running
    s = oracle.exchange_step([10, 5, 50, 1], [big] * R, order, pos, tss=5, ladder_best=big)
    assert list(s["order"]) == [0, 2, 1, 3] and s["exchanges"] == 1 and s["tss"] == 6
    # replica index R-1 moving up resets timeSinceSwap (here: position 2 <-> 3)
    s = oracle.exchange_step([50, 40, 10, 30], [big] * R, order, pos, tss=5, ladder_best=big)
    assert list(s["order"]) == [0, 1, 3, 2] and s["exchanges"] == 1 and s["tss"] == 0
    # equal energies never swap; an improvement in any bestE resets timeSinceSwap
    s = oracle.exchange_step([7, 7, 7, 7], [9, 8, big, big], order, pos, tss=5, ladder_best=10)
    assert s["exchanges"] == 0 and s["tss"] == 0 and s["ladder_best"] == 8
    # kick after the threshold, and again next step (it does not reset the counter)
    s = oracle.exchange_step([7, 7, 7, 7], [big] * R, order, pos, tss=100, ladder_best=1,
                             kick_after=100)
    assert s["kicked"] and list(s["order"]) == [3, 1, 2, 0] and list(s["pos"]) == [3, 1, 2, 0]
    s2 = oracle.exchange_step([7, 7, 7, 7], [big] * R, s["order"], s["pos"], tss=s["tss"],
                              ladder_best=1, kick_after=100)
    assert s2["kicked"] and s2["tss"] == 102 and list(s2["order"]) == [0, 1, 2, 3]


def _brute_force_min(F, D):
    n = F.shape[0]
    best = None
    for p in itertools.permutations(range(n)):
        p = np.array(p)
        e = int((D * F[np.ix_(p, p)]).sum())
        best = e if best is None or e < best else best
    return best


@pytest.mark.parametrize("n,seed", [(6, 1), (7, 2), (8, 3)])
def test_solver_reaches_brute_force_optimum(qpt, oracle, n, seed):
    F, D = qpt.instances.random_instance(300 + seed, n, 0, 20, symmetric=(seed != 2),
                                         zero_diag=(seed != 2))
    opt = _brute_force_min(F.astype(np.int64), D.astype(np.int64))
    betas = qpt.instances.geometric_ladder(8, 0.5, 200.0)
    for mode in (oracle.MODE_EXACT, oracle.MODE_REFFP):
        r = oracle.solve(F, D, betas, 400, seed=seed, mode=mode)
        assert r["best_cost"] == opt
        assert qpt.instances.qap_cost(F, D, r["best_perm"]) == opt
        # every replica's tracked best cost belongs to its recorded permutation
        for k in range(8):
            if r["final_bestE"][k] != np.iinfo(np.int64).max:
                assert qpt.instances.qap_cost(F, D, r["best_perms"][k]) == r["final_bestE"][k]
            assert qpt.instances.qap_cost(F, D, r["final_perm"][k]) == r["final_E"][k]


def test_init_perm_is_identity_plus_transpositions(oracle):
    n, seed = 17, 99
    p = oracle.init_perm(n, seed, 3)
    q = np.arange(n)
    for i in range(n):
        x = oracle.philox([i, 3, 0, 1], [seed & 0xffffffff, seed >> 32])[0]
        j = (x * n) >> 32
        q[i], q[j] = q[j], q[i]
    np.testing.assert_array_equal(p, q)
    assert sorted(p) == list(range(n))


def test_threaded_ladders_equal_sequential(qpt, oracle):
    F, D = qpt.instances.random_instance(41, 20)
    betas = qpt.instances.geometric_ladder(4, 1.0, 50.0)
    a = oracle.solve(F, D, betas, 30, T=5, seed=2, nthreads=1)
    b = oracle.solve(F, D, betas, 30, T=5, seed=2, nthreads=3)
    for k in ("final_perm", "final_E", "final_bestE", "final_pos"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["best_cost"] == b["best_cost"]


def test_golden_trajectory(qpt, oracle):
    with open(os.path.join(GOLDEN, "trajectory_n24.json")) as fh:
        g = json.load(fh)
    F = np.array(g["F"], dtype=np.int32)
    D = np.array(g["D"], dtype=np.int32)
    r = oracle.solve(F, D, np.array(g["beta_ln"]), g["sweeps"], T=g["T"], seed=g["seed"],
                     traj=True)
    np.testing.assert_array_equal(r["traj_E"], np.array(g["traj_E"], dtype=np.int64))
    np.testing.assert_array_equal(r["final_perm"], np.array(g["final_perm"]))
    assert r["best_cost"] == g["best_cost"]
    assert r["accepted"] == g["accepted"] and r["exchanges"] == g["exchanges"]
