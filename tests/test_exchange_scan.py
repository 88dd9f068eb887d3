"""The exchange step's bubble pass as the mirrored exchange step computes it for a single
ladder (qap_pt_xmirror.h, lane k = ladder position k): with prefix-minimum "records" (the
positions whose cost equals the minimum of the costs up to them) the element carried into
step k is the last record at or before k, and step k places the next element at k when
position k + 1 is not a record (the sequential pass's swap, cE < nE), the carried one
otherwise.  This model of that formulation is checked against the oracle's sequential
exchange step (CPU_PIM.cpp:323-344) on random ladders, with ties, as the order, the
exchange count and the reset of the time since the coldest replica moved."""
import numpy as np


def scan_pass(E, order):
    """The lane-parallel pass: new order, exchanges, whether replica R-1 moved up."""
    R = len(order)
    e_pos = np.array([E[r] for r in order])
    pmin = np.minimum.accumulate(e_pos)
    rec = e_pos == pmin
    carry = np.maximum.accumulate(np.where(rec, np.arange(R), -1))
    rcar = np.asarray(order)[carry]
    new = np.empty(R, dtype=np.int64)
    nex, reset = 0, False
    for k in range(R):
        if k < R - 1:
            swap = not rec[k + 1]
            new[k] = order[k + 1] if swap else rcar[k]
            if swap:
                nex += 1
                reset |= order[k + 1] == R - 1
        else:
            new[k] = rcar[k]
    return new, nex, reset


def test_scan_pass_equals_sequential_exchange(oracle):
    rng = np.random.default_rng(17)
    big = np.iinfo(np.int64).max
    for trial in range(400):
        R = int(rng.choice([1, 2, 3, 7, 16, 64]))
        span = int(rng.choice([1, 3, 50, 10**6]))
        E = rng.integers(-span, span + 1, R)
        order = rng.permutation(R)
        pos = np.argsort(order)
        s = oracle.exchange_step(E, [big] * R, order, pos, tss=5, ladder_best=big,
                                 kick_after=10**9)
        new, nex, reset = scan_pass(E, order)
        np.testing.assert_array_equal(s["order"], new, err_msg=f"trial {trial}")
        assert s["exchanges"] == nex, trial
        assert (s["tss"] == 0) == reset, trial
