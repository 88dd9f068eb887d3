"""GPU checks at the benchmark configurations' full sizes.

The bit-exact trajectory cases elsewhere run at sizes the oracle finishes in seconds.  Here
the 512-instance batch (512 instances, n uniform in [30,150], 32 replicas each: 16384
workgroups over four kernel buckets) and the n = 1024 large dense instance with all 64
replicas run at their real shapes, and are checked through properties that hold at any
size: every incrementally kept force matrix and cost equals a from-scratch rebuild on the
device, every reported cost equals the host's recomputation of the QAP objective for the
reported permutation, the attempt counts are exact, and -- for three instances of the batch
chosen by size -- the whole replica state inside the full pool equals the oracle's run of
that instance alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SWEEPS = 20


def _is_perm(p, n):
    return np.array_equal(np.sort(np.asarray(p)), np.arange(n))


def test_batch512_full_size(qpt, oracle, gpu):
    I = qpt.instances
    ladder = I.geometric_ladder(32, 100.0, 10000.0)
    pairs = I.batch_c3()
    insts = [qpt.Instance(F=F, D=D, beta_ln=ladder, sweeps=SWEEPS, seed=7000 + k)
             for k, (F, D) in enumerate(pairs)]
    assert len(insts) == 512
    print(f"batch: {len(insts)} instances x 32 replicas, {SWEEPS} sweeps", flush=True)
    with qpt.Pool(insts) as pool:
        shapes = {pool.shape(k) for k in range(len(insts))}
        pool.run(-1)
        fm, cm = pool.verify()
        res = [pool.result(k) for k in range(len(insts))]
        ns = np.array([x.n for x in insts])
        order = np.argsort(ns, kind="stable")
        picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
        states = {k: pool.state(k) for k in picks}
        # two replicas' states of every instance for the host recomputation
        some = {k: pool.state(k) for k in range(0, len(insts), 8)}
    print(f"shapes in the pool: {sorted(shapes)}", flush=True)
    assert any(",lazy," in s for s in shapes)  # the full-chip placement runs
    assert (fm, cm) == (0, 0), f"force/cost mismatches: {fm}/{cm}"
    for k, (x, r) in enumerate(zip(insts, res)):
        assert r["attempts"] == SWEEPS * x.n * 32, k
        assert r["sweeps_done"] == SWEEPS, k
        assert _is_perm(r["best_perm"], x.n), k
        assert r["best_cost"] == I.qap_cost(x.F, x.D, r["best_perm"]), k
    for k, st in some.items():
        x = insts[k]
        for rep in (0, 31):
            assert _is_perm(st["perm"][rep], x.n)
            assert st["E"][rep] == I.qap_cost(x.F, x.D, st["perm"][rep]), (k, rep)
            assert st["bestE"][rep] == I.qap_cost(x.F, x.D, st["best_perms"][rep]), (k, rep)
    # bit for bit against the oracle: the smallest, the median and the largest instance
    for k in picks:
        x, st = insts[k], states[k]
        print(f"oracle: instance {k}, n={x.n}", flush=True)
        ref = oracle.solve(x.F, x.D, x.beta_ln, SWEEPS, seed=x.seed)
        np.testing.assert_array_equal(st["perm"], ref["final_perm"])
        np.testing.assert_array_equal(st["E"], ref["final_E"])
        np.testing.assert_array_equal(st["bestE"], ref["final_bestE"])
        np.testing.assert_array_equal(st["pos"], ref["final_pos"])
        np.testing.assert_array_equal(st["accepted"], ref["final_accepted"])
        assert res[k]["best_cost"] == ref["best_cost"]
        assert res[k]["exchanges"] == ref["exchanges"] and res[k]["kicks"] == ref["kicks"]


def test_large1024_all_replicas(qpt, gpu):
    I = qpt.instances
    F, D = I.large_c5(1024)
    betas = I.geometric_ladder(64, 300.0, 30000.0)
    inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=2, seed=2024)
    print("n=1024, 64 replicas, 2 sweeps", flush=True)
    with qpt.Pool([inst]) as pool:
        shape = pool.shape(0)
        pool.run(-1)
        fm, cm = pool.verify()
        r = pool.result(0)
        st = pool.state(0)
    assert ",lazy," in shape, shape
    assert (fm, cm) == (0, 0), f"force/cost mismatches: {fm}/{cm}"
    assert r["attempts"] == 2 * 1024 * 64
    assert r["flush_cells"] > 0 and r["flush_cells"] % (1024 * 1024) == 0
    assert r["best_cost"] == I.qap_cost(F, D, r["best_perm"])
    for rep in range(64):
        assert _is_perm(st["perm"][rep], 1024)
        assert st["E"][rep] == I.qap_cost(F, D, st["perm"][rep]), rep
        assert st["bestE"][rep] == I.qap_cost(F, D, st["best_perms"][rep]), rep
