"""Pins of the oracle's one-shot parallel-proposal step (GPU-CUPIM/perl-gpu/src/lib/qap/
qap.cu:1037-1460): every location proposes from one snapshot, claims resolve conflicts,
the disjoint winners' rank-2 updates are summed and the cost is recomputed.

Each check uses something the step does not compute itself: swap deltas from
from-scratch costs, force matrices rebuilt from the permutation, the conflict rule
restated as pairwise comparisons, and proposal frequencies against the Gibbs law.
"""
import itertools

import numpy as np
import pytest


def _brute_deltas(oracle, F, D, p):
    """M[i][j] = E(p) - E(p with the facilities at i and j swapped), from full costs."""
    n = len(p)
    E = oracle.energy(F, D, p)
    M = np.zeros((n, n), np.int64)
    for i in range(n):
        for j in range(n):
            q = p.copy()
            q[i], q[j] = q[j], q[i]
            M[i, j] = E - oracle.energy(F, D, q)
    return M


def _start(qpt, oracle, seed, n, symmetric=True, zero_diag=True, hi=9):
    F, D = qpt.instances.random_instance(seed, n, 0, hi, symmetric=symmetric,
                                         zero_diag=zero_diag)
    p = np.random.default_rng(seed).permutation(n).astype(np.int32)
    return F, D, p, oracle.force_matrix(F, D, p), oracle.energy(F, D, p)


@pytest.mark.parametrize("symmetric,zero_diag", [(True, True), (False, False), (False, True)])
def test_step_keeps_force_matrix_and_cost_exact(qpt, oracle, symmetric, zero_diag):
    """After every step the incrementally updated force matrix equals a rebuild from the
    new permutation, the cost equals the full sum, and the permutation moved by exactly
    the applied (disjoint) swaps."""
    F, D, p, I, E = _start(qpt, oracle, 51, 14, symmetric, zero_diag)
    b2 = oracle.beta2(0.05)
    total = 0
    for step in range(12):
        m, tgt, won, p2, I2, E2 = oracle.one_shot_step(F, D, p, I, E, np.float32(b2),
                                                       20.0 / b2, 77, step, 3)
        np.testing.assert_array_equal(I2, oracle.force_matrix(F, D, p2))
        assert E2 == oracle.energy(F, D, p2)
        assert sorted(p2) == list(range(14))
        q = p.copy()
        touched = set()
        for i in np.flatnonzero(won):
            j = int(tgt[i])
            assert i not in touched and j not in touched and i != j
            touched |= {int(i), j}
            q[i], q[j] = p[j], p[i]
        np.testing.assert_array_equal(p2, q)
        assert m == int(won.sum())
        total += m
        p, I, E = p2, I2, E2
    assert total > 0


def test_conflict_rule_largest_delta_then_smallest_location(qpt, oracle):
    """A proposal is applied exactly when no other proposal claiming one of its two
    locations has a larger delta, or the same delta from a smaller location."""
    F, D, p, I, E = _start(qpt, oracle, 52, 12)
    b2 = oracle.beta2(0.002)  # hot: most locations propose a move
    seen_lost = 0
    for step in range(20):
        M = _brute_deltas(oracle, F, D, p)
        m, tgt, won, p2, I2, E2 = oracle.one_shot_step(F, D, p, I, E, np.float32(b2),
                                                       20.0 / b2, 5, step, 0)
        movers = [i for i in range(12) if tgt[i] != i]
        for i in movers:
            locs = {i, int(tgt[i])}
            beaten = False
            for k in movers:
                if k == i or not ({k, int(tgt[k])} & locs):
                    continue
                dk, di = M[k, tgt[k]], M[i, tgt[i]]
                if dk > di or (dk == di and k < i):
                    beaten = True
            assert won[i] == (not beaten), (step, i)
            seen_lost += beaten
        for i in range(12):
            if tgt[i] == i:
                assert not won[i]
        p, I, E = p2, I2, E2
    assert seen_lost > 0


def test_proposals_follow_the_heat_bath_law(qpt, oracle):
    """On a frozen state, location i's proposal over many steps is drawn with
    probability 2^(beta2 * M[i][j]) / Z, M from full costs (stay included, j = i)."""
    F, D, p, I, E = _start(qpt, oracle, 53, 6)
    M = _brute_deltas(oracle, F, D, p).astype(np.float64)
    b2 = oracle.beta2(0.01)
    N = 3000
    counts = np.zeros((6, 6))
    for step in range(N):
        _, tgt, _, _, _, _ = oracle.one_shot_step(F, D, p, I, E, np.float32(b2), 20.0 / b2,
                                                  9, step, 1)
        counts[np.arange(6), tgt] += 1
    for i in range(6):
        z = 2.0 ** (b2 * (M[i] - M[i].max()))
        prob = z / z.sum()
        sd = np.sqrt(prob * (1 - prob) / N)
        assert np.all(np.abs(counts[i] / N - prob) <= 5 * sd + 1e-3), (i, counts[i] / N, prob)


def test_cold_limit_proposes_a_best_swap(qpt, oracle):
    F, D, p, I, E = _start(qpt, oracle, 54, 9)
    b2 = oracle.beta2(1e3)
    for step in range(6):
        M = _brute_deltas(oracle, F, D, p)
        m, tgt, won, p2, I2, E2 = oracle.one_shot_step(F, D, p, I, E, np.float32(b2),
                                                       20.0 / b2, 11, step, 2)
        for i in range(9):
            assert M[i, tgt[i]] == M[i].max()
        for i in np.flatnonzero(won):
            assert M[i, tgt[i]] >= 0
        if m == 1:  # one swap: the cost moves by exactly its delta
            i = int(np.flatnonzero(won)[0])
            assert E2 == E - M[i, tgt[i]]
        # several disjoint swaps interact (qap.cu rescores for that reason): the cost
        # after the step is the full cost of the new permutation, whatever its sign
        assert E2 == oracle.energy(F, D, p2)
        p, I, E = p2, I2, E2


def test_two_locations_mutual_proposal_applies_once(oracle):
    """n = 2: both locations propose the same improving swap; the smaller location's
    claim wins both locations and the swap is applied once."""
    F = np.array([[9, 0], [0, 1]], np.int32)
    D = np.array([[1, 0], [0, 5]], np.int32)
    p = np.array([1, 0], np.int32)
    I = oracle.force_matrix(F, D, p)
    E = oracle.energy(F, D, p)
    assert E == 46
    b2 = oracle.beta2(1e3)
    m, tgt, won, p2, I2, E2 = oracle.one_shot_step(F, D, p, I, E, np.float32(b2), 20.0 / b2,
                                                   1, 0, 0)
    np.testing.assert_array_equal(tgt, [1, 0])
    np.testing.assert_array_equal(won, [True, False])
    assert m == 1 and E2 == 14
    np.testing.assert_array_equal(p2, [0, 1])


def _brute_force_min(F, D):
    n = F.shape[0]
    best = None
    for q in itertools.permutations(range(n)):
        q = np.array(q)
        e = int((D * F[np.ix_(q, q)]).sum())
        best = e if best is None or e < best else best
    return best


@pytest.mark.parametrize("n,seed", [(6, 1), (7, 2)])
def test_one_shot_solver_reaches_brute_force_optimum(qpt, oracle, n, seed):
    F, D = qpt.instances.random_instance(310 + seed, n, 0, 20, symmetric=(seed != 2),
                                         zero_diag=(seed != 2))
    opt = _brute_force_min(F.astype(np.int64), D.astype(np.int64))
    betas = qpt.instances.geometric_ladder(8, 0.5, 200.0)
    r = oracle.solve(F, D, betas, 400, seed=seed, one_shot=True)
    assert r["best_cost"] == opt
    assert qpt.instances.qap_cost(F, D, r["best_perm"]) == opt
    for k in range(8):
        if r["final_bestE"][k] != np.iinfo(np.int64).max:
            assert qpt.instances.qap_cost(F, D, r["best_perms"][k]) == r["final_bestE"][k]
        assert qpt.instances.qap_cost(F, D, r["final_perm"][k]) == r["final_E"][k]
