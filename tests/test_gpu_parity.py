"""GPU parity: the HIP sweep/exchange kernels against the CPU oracle, bit for bit.

Every case runs the same seeded instance, ladder and Philox stream through the
C-ABI (libqap_pt.so) and through oracle/qap_oracle.c, then compares the complete
replica state: permutations, exact costs, per-replica bests and their permutations,
ladder positions, accepted-move counts, exchange/kick counts, and the best result.
The incrementally tracked force matrices and costs are also rebuilt from scratch on
the device (qap_pt_pool_verify) and must agree exactly.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_both(qpt, oracle, F, D, betas, sweeps, T=1, spe=1, seed=7, kick_after=100, goal=None,
              per_replica_init=False, env=None, init_perm=None, one_shot=False):
    env = env or {}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=sweeps, T=T, sweeps_per_exchange=spe,
                            seed=seed, kick_after=kick_after, goal=goal,
                            per_replica_init=per_replica_init, init_perm=init_perm,
                            one_shot=one_shot)
        with qpt.Pool([inst]) as pool:
            pool.run(-1)
            res = pool.result(0)
            st = pool.state(0)
            fm, cm = pool.verify()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    ref = oracle.solve(F, D, betas, sweeps, T=T, spe=spe, seed=seed, kick_after=kick_after,
                       goal=goal, per_replica_init=per_replica_init, init_perm=init_perm,
                       one_shot=one_shot)
    return res, st, (fm, cm), ref


def _assert_same(res, st, ver, ref):
    assert ver == (0, 0), f"force/cost mismatches after the run: {ver}"
    np.testing.assert_array_equal(st["E"], ref["final_E"])
    np.testing.assert_array_equal(st["perm"], ref["final_perm"])
    np.testing.assert_array_equal(st["bestE"], ref["final_bestE"])
    np.testing.assert_array_equal(st["pos"], ref["final_pos"])
    np.testing.assert_array_equal(st["accepted"], ref["final_accepted"])
    np.testing.assert_array_equal(st["best_perms"], ref["best_perms"])
    assert res["best_cost"] == ref["best_cost"]
    np.testing.assert_array_equal(res["best_perm"], ref["best_perm"])
    assert res["accepted"] == ref["accepted"]
    assert res["exchanges"] == ref["exchanges"]
    assert res["kicks"] == ref["kicks"]
    assert res["attempts"] == ref["attempts"]


def _sparse_asym(seed, n, density):
    """Asymmetric flow with nonzero diagonal and few entries per row (most swaps change
    a handful of flow columns: the short column walk with both update terms)."""
    g = np.random.default_rng(seed)
    F = np.where(g.random((n, n)) < density, g.integers(1, 40, (n, n)), 0).astype(np.int32)
    D = g.integers(0, 30, (n, n)).astype(np.int32)
    return F, D


def _wide(seed, n, wide_d, symmetric=True, density=0.15, big=4_300_000):
    """One matrix 0/1 at the given density, the other with entries up to `big`, beyond
    the 24-bit multiplier's range (2 max|D| or 4 max|F| >= 2^23): every update takes
    the full 32-bit products.  The defaults at n = 96 start near a cost of 3e9, beyond
    int32, with the force and delta bounds still inside the host's int32 check."""
    g = np.random.default_rng(seed)
    B = (g.random((n, n)) < density).astype(np.int32)
    W = g.integers(0, big, (n, n)).astype(np.int32)
    if symmetric:
        B = np.triu(B, 1) + np.triu(B, 1).T
        W = np.triu(W, 1) + np.triu(W, 1).T
    return (B, W) if wide_d else (W, B)


def _twins(seed, n, blocks=((0, 8, 4), (8, 16, 9), (16, 24, 0)), pairs=((30, 31, 3), (33, 40, 6))):
    """Symmetric zero-diagonal flow with planted twin sets (flow rows equal outside the
    set's own columns): blocks [lo, hi) with pair value f and disjoint twin pairs (a, b, f).
    A swap of twins changes two columns of the force matrix (one block with f = 0 has
    identical rows: its swaps change nothing)."""
    g = np.random.default_rng(seed)
    F = np.triu(g.integers(0, 6, (n, n)), 1)
    F = (F + F.T).astype(np.int32)
    for lo, hi, f in blocks:
        for a in range(lo + 1, hi):
            for x in range(n):
                if not lo <= x < hi:
                    F[a, x] = F[x, a] = F[lo, x]
        for a in range(lo, hi):
            for b in range(lo, hi):
                F[a, b] = f if a != b else 0
    for a, b, f in pairs:
        for x in range(n):
            if x not in (a, b):
                F[b, x] = F[x, b] = F[a, x]
        F[a, b] = F[b, a] = f
    D = np.triu(g.integers(0, 20, (n, n)), 1)
    return F, (D + D.T).astype(np.int32)


def _wide_costs(seed, n):
    """Symmetric zero-diagonal flow in [0, 2^16) and distances in [0, 40]: the force bound
    fits int32 but the cost bound max|F| sum|D| passes 2^31."""
    g = np.random.default_rng(seed)
    F = np.triu(g.integers(0, 1 << 16, (n, n)), 1)
    D = np.triu(g.integers(0, 41, (n, n)), 1)
    return (F + F.T).astype(np.int32), (D + D.T).astype(np.int32)


def _wide64(seed, n, symmetric=True, zero_diag=True, big=1 << 27):
    """Flow entries up to 2^27 and distances up to 63: force bounds beyond int32, so the
    instance runs with int64 force matrices; costs reach ~1e13."""
    g = np.random.default_rng(seed)
    F = g.integers(0, big, (n, n), dtype=np.int64)
    D = g.integers(0, 64, (n, n), dtype=np.int64)
    if symmetric:
        F = np.triu(F, 1) + np.triu(F, 1).T
        D = np.triu(D, 1) + np.triu(D, 1).T
    if zero_diag:
        np.fill_diagonal(F, 0)
        np.fill_diagonal(D, 0)
    return F.astype(np.int32), D.astype(np.int32)


def _wide_scale(n):
    """beta scale that puts the int64 instances' deltas (~4 * 63 * 2^27 * n) at ~20 / beta"""
    return 20.0 / (4 * 63 * (1 << 27) * n) * max(1.0, n / 16.0)


CASES = {
    # name: (instance builder, R, sweeps, extra kwargs)
    "sym_n16_R4": (lambda q: q.instances.random_instance(11, 16), 4, 40, {}),
    "c1_n64_R16": (lambda q: q.instances.flat_c1(), 16, 12, {}),
    "sym_n100_R8": (lambda q: q.instances.mixed_instance(5, 100, True), 8, 6, {}),
    "sparse_n150_R4": (lambda q: q.instances.mixed_instance(6, 150, False), 4, 4, {}),
    "asym_diag_n24_T2_spe2": (lambda q: q.instances.random_instance(12, 24, symmetric=False,
                                                                   zero_diag=False), 6, 20,
                              {"T": 2, "spe": 2}),
    "asym_n70_R5": (lambda q: q.instances.random_instance(13, 70, 0, 50, symmetric=False), 5, 6,
                    {}),
    "hbm_n40_R6": (lambda q: q.instances.random_instance(14, 40), 6, 10,
                   {"env": {"QAP_PT_IMODE": "2"}}),
    "hbm_asym_n37": (lambda q: q.instances.random_instance(19, 37, symmetric=False,
                                                          zero_diag=False), 4, 8,
                     {"env": {"QAP_PT_IMODE": "2"}}),
    "lds16_n50_R6": (lambda q: q.instances.random_instance(20, 50), 6, 10,
                     {"env": {"QAP_PT_IMODE": "1"}}),
    "lds16_asym_n45": (lambda q: q.instances.random_instance(23, 45, symmetric=False,
                                                            zero_diag=False), 4, 8,
                       {"env": {"QAP_PT_IMODE": "1"}}),
    "lds16_grey_n256_R4": (lambda q: q.instances.grey_c2(), 4, 3, {}),
    "sparse_n90_64x2": (lambda q: q.instances.mixed_instance(24, 90, False), 5, 8,
                        {"env": {"QAP_PT_SHAPE": "64,2"}}),
    "sparse_n120_128x1_hbm": (lambda q: q.instances.mixed_instance(25, 120, False), 4, 6,
                              {"env": {"QAP_PT_SHAPE": "128,1", "QAP_PT_IMODE": "2"}}),
    "dense_n77_64x2": (lambda q: q.instances.mixed_instance(26, 77, True), 5, 8,
                       {"env": {"QAP_PT_SHAPE": "64,2"}}),
    "n1_trivial": (lambda q: q.instances.random_instance(27, 1), 3, 5, {}),
    "n2_R2": (lambda q: q.instances.random_instance(28, 2), 2, 20, {}),
    "per_replica_init_n33_T3": (lambda q: q.instances.random_instance(15, 33), 5, 10,
                                {"T": 3, "per_replica_init": True}),
    "kick_often_n20": (lambda q: q.instances.random_instance(16, 20), 4, 60, {"kick_after": 2}),
    # exchange step: many ladders (one thread per ladder) and a ladder longer than a wave
    "ladders_T12_kick": (lambda q: q.instances.random_instance(32, 24), 4, 30,
                         {"T": 12, "kick_after": 3}),
    "long_ladder_R70": (lambda q: q.instances.random_instance(33, 16), 70, 8, {"kick_after": 2}),
    "multiwave_n64_256x1": (lambda q: q.instances.flat_c1(), 4, 6,
                            {"env": {"QAP_PT_SHAPE": "256,1"}}),
    "multiwave_n64_128x1": (lambda q: q.instances.flat_c1(), 4, 6,
                            {"env": {"QAP_PT_SHAPE": "128,1"}}),
    "cpl4_n200_64x4": (lambda q: q.instances.mixed_instance(17, 200, True), 3, 2,
                       {"env": {"QAP_PT_SHAPE": "64,4"}}),
    "grey_n256_R4_hbm": (lambda q: q.instances.grey_c2(), 4, 2, {"env": {"QAP_PT_IMODE": "2"}}),
    "n300_256x2_hbm": (lambda q: q.instances.mixed_instance(18, 300, True), 2, 1, {}),
    # force matrix in HBM with rank-2 updates batched in LDS (symmetric instances)
    "lazy_n40_R4": (lambda q: q.instances.random_instance(29, 40), 4, 12,
                    {"env": {"QAP_PT_IMODE": "3"}}),
    # flush helpers: three helper workgroups per replica share every flush (rows split
    # four ways, pending updates read from global memory)
    "lazy_n40_R4_3helpers": (lambda q: q.instances.random_instance(29, 40), 4, 12,
                             {"env": {"QAP_PT_IMODE": "3", "QAP_PT_FLUSH_HELPERS": "3"}}),
    "lazy_n200_T2_2helpers": (lambda q: q.instances.mixed_instance(31, 200, True), 3, 4,
                              {"T": 2, "kick_after": 1,
                               "env": {"QAP_PT_IMODE": "3", "QAP_PT_FLUSH_HELPERS": "2"}}),
    "lazy_c5_n512_1helper": (lambda q: q.instances.large_c5(512), 2, 1,
                             {"env": {"QAP_PT_FLUSH_HELPERS": "1"}}),
    "lazy_sparse_n90_64x2": (lambda q: q.instances.mixed_instance(30, 90, False), 4, 8,
                             {"env": {"QAP_PT_IMODE": "3", "QAP_PT_SHAPE": "64,2"}}),
    "lazy_grey_n256_R3": (lambda q: q.instances.grey_c2(), 3, 3, {"env": {"QAP_PT_IMODE": "3"}}),
    "lazy_n200_T2_kick": (lambda q: q.instances.mixed_instance(31, 200, True), 3, 4,
                          {"T": 2, "kick_after": 1, "env": {"QAP_PT_IMODE": "3"}}),
    "lazy_c5_n512": (lambda q: q.instances.large_c5(512), 2, 1, {}),
    # eight waves per replica
    "lazy_n200_512x1": (lambda q: q.instances.mixed_instance(33, 200, True), 3, 3,
                        {"env": {"QAP_PT_IMODE": "3", "QAP_PT_SHAPE": "512,1"}}),
    "lazy_c5_n512_512x1": (lambda q: q.instances.large_c5(512), 2, 1,
                           {"env": {"QAP_PT_SHAPE": "512,1"}}),
    "lazy_n700_512x2": (lambda q: q.instances.random_instance(34, 700), 2, 1,
                        {"env": {"QAP_PT_IMODE": "3", "QAP_PT_SHAPE": "512,2"}}),
    # sparse asymmetric flow: short column walk (u2, v2 terms) in LDS and in HBM
    "sparse_asym_n60": (lambda q: _sparse_asym(40, 60, 0.04), 4, 10, {}),
    "sparse_asym_n60_hbm": (lambda q: _sparse_asym(41, 60, 0.04), 4, 10,
                            {"env": {"QAP_PT_IMODE": "2"}}),
    "sparse_asym_n200_lds16": (lambda q: _sparse_asym(42, 200, 0.03), 3, 3, {}),
    # planted twin facilities: the two-column update (uint16 LDS and HBM placements; the
    # int32 LDS kernel takes the general update on the same instances)
    "twins_n48": (lambda q: _twins(60, 48), 4, 12, {}),
    "twins_n48_hbm": (lambda q: _twins(61, 48), 4, 12, {"env": {"QAP_PT_IMODE": "2"}}),
    "twins_n48_lds16": (lambda q: _twins(62, 48), 4, 12, {"env": {"QAP_PT_IMODE": "1"}}),
    # twin updates of the uint16 placement land one update late (the next site's deltas
    # carry their terms): longer runs, the one-wave shape, and partner sites next in line
    "twins_n48_lds16_40sw": (lambda q: _twins(66, 48), 4, 40, {"env": {"QAP_PT_IMODE": "1"}}),
    "twins_n48_64x4_lds16": (lambda q: _twins(67, 48), 4, 30,
                             {"env": {"QAP_PT_IMODE": "1", "QAP_PT_SHAPE": "64,4"}}),
    "twins_n9_lds16_kick": (lambda q: _twins(68, 9, blocks=((0, 5, 3),), pairs=()), 5, 60,
                            {"kick_after": 2, "env": {"QAP_PT_IMODE": "1"}}),
    "grey_n256_R8_20sw": (lambda q: q.instances.grey_c2(), 8, 20, {}),
    "grey_n256_R4_64x4": (lambda q: q.instances.grey_c2(), 4, 4, {"env": {"QAP_PT_SHAPE": "64,4"}}),
    "grey_n256_R4_128x2": (lambda q: q.instances.grey_c2(), 4, 4,
                           {"env": {"QAP_PT_SHAPE": "128,2"}}),
    "twins_n48_64x1_T2_lds16": (lambda q: _twins(63, 48), 3, 12,
                                {"T": 2, "env": {"QAP_PT_SHAPE": "64,1", "QAP_PT_IMODE": "1"}}),
    "twins_n150_128x2_lds16": (lambda q: _twins(64, 150, blocks=((0, 40, 2), (40, 60, 0)),
                                                pairs=((100, 149, 5),)), 3, 4,
                               {"env": {"QAP_PT_SHAPE": "128,2", "QAP_PT_IMODE": "1"}}),
    "twins_n150_256x1_hbm": (lambda q: _twins(65, 150, blocks=((10, 50, 3),),
                                              pairs=((100, 101, 1), (7, 149, 2))), 3, 3,
                             {"env": {"QAP_PT_SHAPE": "256,1", "QAP_PT_IMODE": "2"}}),
    # entries beyond the 24-bit multiplier, costs beyond int32 (beta scaled to the deltas)
    "wide_d_n96": (lambda q: _wide(50, 96, True), 4, 6, {"beta_scale": 1e-5}),
    "wide_d_asym_n96_hbm": (lambda q: _wide(51, 96, True, symmetric=False), 4, 6,
                            {"beta_scale": 1e-5, "env": {"QAP_PT_IMODE": "2"}}),
    "wide_f_n96_64x2": (lambda q: _wide(52, 96, False, density=0.2, big=2_500_000), 4, 6,
                        {"beta_scale": 1e-5, "env": {"QAP_PT_SHAPE": "64,2"}}),
    "wide_f_asym_n64": (lambda q: _wide(54, 64, False, symmetric=False, density=0.2,
                                        big=2_500_000), 4, 8, {"beta_scale": 1e-5}),
    # one-shot parallel-proposal steps (qap.cu:1037-1609): 64 / 128 / 256 threads
    "os_sym_n16_R4": (lambda q: q.instances.random_instance(11, 16), 4, 40, {"one_shot": True}),
    "os_c1_n64_R16": (lambda q: q.instances.flat_c1(), 16, 12, {"one_shot": True}),
    "os_n37_T2_spe3": (lambda q: q.instances.random_instance(35, 37), 4, 12,
                       {"one_shot": True, "T": 2, "spe": 3}),
    "os_sym_n100_R8": (lambda q: q.instances.mixed_instance(5, 100, True), 8, 6,
                       {"one_shot": True}),
    "os_sparse_n150_R4": (lambda q: q.instances.mixed_instance(6, 150, False), 4, 4,
                          {"one_shot": True}),
    "os_asym_diag_n24_T2": (lambda q: q.instances.random_instance(12, 24, symmetric=False,
                                                                 zero_diag=False), 6, 20,
                            {"one_shot": True, "T": 2, "spe": 2}),
    "os_asym_n70_R5": (lambda q: q.instances.random_instance(13, 70, 0, 50, symmetric=False),
                       5, 6, {"one_shot": True}),
    "os_asym_n200_R2": (lambda q: q.instances.random_instance(36, 200, 0, 30, symmetric=False),
                        2, 2, {"one_shot": True}),
    "os_grey_n256_R3": (lambda q: q.instances.grey_c2(), 3, 3, {"one_shot": True}),
    "os_n256_hot_R2": (lambda q: q.instances.mixed_instance(37, 256, True), 2, 2,
                       {"one_shot": True}),
    "os_grey_256x4": (lambda q: q.instances.grey_c2(), 2, 3,
                      {"one_shot": True, "env": {"QAP_PT_SHAPE": "256,4"}}),
    "os_asym_n100_256x2": (lambda q: q.instances.random_instance(38, 100, 0, 30, symmetric=False),
                           3, 4, {"one_shot": True, "env": {"QAP_PT_SHAPE": "256,2"}}),
    "os_n50_256x1": (lambda q: q.instances.random_instance(39, 50), 3, 6,
                     {"one_shot": True, "env": {"QAP_PT_SHAPE": "256,1"}}),
    # entries beyond the int32 force bound: int64 force matrices in HBM (hbm64 placement)
    "wide64_sym_n40": (lambda q: _wide64(80, 40), 4, 10,
                       {"beta_scale": _wide_scale(40), "expect_shape": "64x1,hbm64"}),
    "wide64_asym_diag_n37_T2": (lambda q: _wide64(81, 37, symmetric=False, zero_diag=False), 3, 8,
                                {"T": 2, "beta_scale": _wide_scale(37), "kick_after": 2,
                                 "expect_shape": "64x1,hbm64,general"}),
    "wide64_n100_128x1": (lambda q: _wide64(82, 100), 3, 4,
                          {"beta_scale": _wide_scale(100), "expect_shape": "128x1,hbm64"}),
    "wide64_n300_256x2": (lambda q: _wide64(83, 300), 2, 1,
                          {"beta_scale": _wide_scale(300), "expect_shape": "256x2,hbm64"}),
    "wide64_n130_256x1_spe2": (lambda q: _wide64(84, 130, symmetric=False), 2, 4,
                               {"spe": 2, "beta_scale": _wide_scale(130),
                                "expect_shape": "256x1,hbm64,general"}),
    # persistent launches of several waves: 2, 4 and 8 candidates per lane in the pick,
    # int32 and uint16 placements, symmetric and general updates
    "lat_asym_n90_128x1": (lambda q: q.instances.random_instance(70, 90, 0, 30, symmetric=False),
                           4, 8, {"env": {"QAP_PT_SHAPE": "128,1"}}),
    "lat_sym_n200_128x2": (lambda q: q.instances.random_instance(71, 200), 3, 4,
                           {"env": {"QAP_PT_SHAPE": "128,2"}}),
    "lat_asym_diag_n240_256x2_lds16": (lambda q: q.instances.random_instance(
        72, 240, 0, 12, symmetric=False, zero_diag=False), 3, 3,
        {"env": {"QAP_PT_SHAPE": "256,2", "QAP_PT_IMODE": "1"}}),
    "lat_sparse_n150_T2_kick": (lambda q: q.instances.mixed_instance(73, 150, False), 4, 12,
                                {"T": 2, "kick_after": 2}),
    # the benchmark configurations at their own sizes, shapes and ladders
    # n = 1024 large dense: lazy placement, eight waves x 2 candidates, 24 pending updates
    "lazy_c5_n1024_R2": (lambda q: q.instances.large_c5(1024), 2, 1,
                         {"expect_shape": "512x2,lazy,sym"}),
    # n = 64 flat: 16 replicas on T 1..100, a thousand exchange intervals
    "c1_n64_R16_1000sweeps": (lambda q: q.instances.flat_c1(), 16, 1000,
                              {"ladder": (1.0, 100.0), "expect_shape": "64x1,lds32,sym"}),
    # grey pattern n = 256: all 64 replicas on T 10..1000 (the all-sparse uint16 build)
    "grey_c2_n256_R64": (lambda q: q.instances.grey_c2(), 64, 3,
                         {"ladder": (10.0, 1000.0),
                          "expect_shape": "256x1,lds16,sym,sparse-flow"}),
    # ... for 20 exchange steps in one persistent launch: the mirrored step's parallel
    # bubble pass over a full wave of ladder positions
    "grey_c2_n256_R64_20sw": (lambda q: q.instances.grey_c2(), 64, 20,
                              {"ladder": (10.0, 1000.0),
                               "expect_shape": "256x1,lds16,sym,sparse-flow"}),
    # a ladder of one replica (the pass with no step)
    "single_replica_n20": (lambda q: q.instances.random_instance(91, 20), 1, 30, {}),
    # several sites per round (sweep_spec_kernel, 16 by default, the latency-bound default
    # for symmetric n <= 64 in int32 LDS) and one site at a time, on mostly rejected updates
    # (the n = 64 flat instance on T 1..100) and on mostly accepted ones (a hot ladder)
    "c1_n64_R16_spec16": (lambda q: q.instances.flat_c1(), 16, 200,
                          {"ladder": (1.0, 100.0), "expect_shape": "64x1,lds32,sym,spec16"}),
    "c1_n64_R16_spec8": (lambda q: q.instances.flat_c1(), 16, 100,
                         {"ladder": (1.0, 100.0), "env": {"QAP_PT_SPEC": "8"},
                          "expect_shape": "64x1,lds32,sym,spec8"}),
    "c1_n64_R16_spec4": (lambda q: q.instances.flat_c1(), 16, 100,
                         {"ladder": (1.0, 100.0), "env": {"QAP_PT_SPEC": "4"},
                          "expect_shape": "64x1,lds32,sym,spec4"}),
    "c1_n64_R16_spec2": (lambda q: q.instances.flat_c1(), 16, 100,
                         {"ladder": (1.0, 100.0), "env": {"QAP_PT_SPEC": "2"},
                          "expect_shape": "64x1,lds32,sym,spec2"}),
    "c1_n64_R16_one_site": (lambda q: q.instances.flat_c1(), 16, 100,
                            {"ladder": (1.0, 100.0), "env": {"QAP_PT_SPEC": "0"},
                             "expect_shape": "64x1,lds32,sym,dense"}),
    "spec16_hot_n50": (lambda q: q.instances.random_instance(80, 50), 8, 30,
                       {"beta_scale": 0.01, "expect_shape": "64x1,lds32,sym,spec16"}),
    "spec8_hot_n50": (lambda q: q.instances.random_instance(80, 50), 8, 30,
                      {"beta_scale": 0.01, "env": {"QAP_PT_SPEC": "8"},
                       "expect_shape": "64x1,lds32,sym,spec8"}),
    "spec16_n2": (lambda q: q.instances.random_instance(81, 2), 4, 20,
                  {"expect_shape": "64x1,lds32,sym,spec16"}),
    "spec4_n2": (lambda q: q.instances.random_instance(81, 2), 4, 20,
                 {"env": {"QAP_PT_SPEC": "4"}, "expect_shape": "64x1,lds32,sym,spec4"}),
    "spec2_n1": (lambda q: q.instances.random_instance(82, 1), 3, 5,
                 {"env": {"QAP_PT_SPEC": "2"}}),
    "spec4_n63_T3_kick": (lambda q: q.instances.random_instance(83, 63), 5, 40,
                          {"T": 3, "kick_after": 2, "env": {"QAP_PT_SPEC": "4"}}),
    "spec8_n63_T3_kick": (lambda q: q.instances.random_instance(83, 63), 5, 40,
                          {"T": 3, "kick_after": 2, "env": {"QAP_PT_SPEC": "8"}}),
    "spec16_n63_T3_kick": (lambda q: q.instances.random_instance(83, 63), 5, 40,
                           {"T": 3, "kick_after": 2}),
    "spec8_hot_n33_kick": (lambda q: q.instances.random_instance(85, 33), 6, 30,
                           {"beta_scale": 0.01, "kick_after": 2, "env": {"QAP_PT_SPEC": "8"},
                            "expect_shape": "64x1,lds32,sym,spec8"}),
    "spec16_n5": (lambda q: q.instances.random_instance(86, 5), 4, 20, {}),
    "spec16_hot_n21_kick": (lambda q: q.instances.random_instance(88, 21), 5, 30,
                            {"beta_scale": 0.01, "kick_after": 2,
                             "expect_shape": "64x1,lds32,sym,spec16"}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_trajectory_bit_exact(qpt, oracle, gpu, name):
    build, R, sweeps, kw = CASES[name]
    kw = dict(kw)
    F, D = build(qpt)
    n = F.shape[0]
    ladder = kw.pop("ladder", None)
    expect_shape = kw.pop("expect_shape", None)
    if ladder is None:
        betas = (qpt.instances.geometric_ladder(R, 1.0, 100.0) / max(1.0, n / 16.0)
                 * kw.pop("beta_scale", 1.0))
    else:  # the workload's own ladder (temperatures t_min .. t_max)
        betas = qpt.instances.geometric_ladder(R, *ladder)
    print(f"case {name}: n={n} R={R} sweeps={sweeps} kw={kw}", flush=True)
    if expect_shape is not None:  # under the case's environment, as the run below
        env = kw.get("env", {})
        old = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            with qpt.Pool([qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=0)]) as pool:
                shape = pool.shape(0)
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        assert shape.startswith(expect_shape), shape
    res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, sweeps, **kw)
    _assert_same(res, st, ver, ref)


def test_flush_cells_count(qpt, oracle, gpu):
    """flush_cells counts the entries the lazy placement's flushes write (n^2 per flush):
    between one flush per 64 accepted moves and one per accepted move, and at most one per
    replica per launch beyond one per 8 moves; in-LDS placements update move by move and
    report none.  Trajectories still equal the oracle's."""
    F, D = qpt.instances.random_instance(29, 40)
    n, R, sweeps = 40, 4, 12
    betas = qpt.instances.geometric_ladder(R, 1.0, 100.0) / (n / 16.0)
    res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, sweeps, env={"QAP_PT_IMODE": "3"})
    _assert_same(res, st, ver, ref)
    acc, fc = res["accepted"], res["flush_cells"]
    assert acc > 0 and fc % (n * n) == 0, (acc, fc)
    flushes = fc // (n * n)
    assert acc / 64 <= flushes <= min(acc, R * sweeps + acc / 8), (acc, flushes)
    res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, sweeps, env={"QAP_PT_IMODE": "0"})
    _assert_same(res, st, ver, ref)
    assert res["flush_cells"] == 0 and res["update_cells"] > 0


def test_goal_stops_like_oracle(qpt, oracle, gpu):
    F, D = qpt.instances.random_instance(21, 30)
    betas = qpt.instances.geometric_ladder(6, 1.0, 50.0)
    free = oracle.solve(F, D, betas, 60, seed=3)
    goal = int(free["best_cost"] * 1.02)
    res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, 60, seed=3, goal=goal)
    assert ref["megas_done"] < 60
    assert res["sweeps_done"] == ref["megas_done"]
    _assert_same(res, st, ver, ref)


def test_goal_stops_like_oracle_one_shot(qpt, oracle, gpu):
    """A goal ends a one-shot run at the same exchange interval as the oracle's."""
    F, D = qpt.instances.random_instance(41, 30)
    betas = qpt.instances.geometric_ladder(6, 1.0, 50.0)
    free = oracle.solve(F, D, betas, 60, seed=3, one_shot=True)
    goal = int(free["best_cost"] * 1.02)
    res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, 60, seed=3, goal=goal,
                                  one_shot=True)
    assert ref["megas_done"] < 60
    assert res["sweeps_done"] == ref["megas_done"]
    _assert_same(res, st, ver, ref)


def test_incremental_runs_equal_one_run_one_shot(qpt, oracle, gpu):
    F, D = qpt.instances.random_instance(42, 48)
    betas = qpt.instances.geometric_ladder(8, 1.0, 100.0) / 3.0
    inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=30, seed=9, one_shot=True)
    with qpt.Pool([inst]) as pool:
        pool.run(10)
        pool.run(20)
        st = pool.state(0)
    ref = oracle.solve(F, D, betas, 30, seed=9, one_shot=True)
    np.testing.assert_array_equal(st["perm"], ref["final_perm"])
    np.testing.assert_array_equal(st["E"], ref["final_E"])
    np.testing.assert_array_equal(st["bestE"], ref["final_bestE"])


def test_incremental_runs_equal_one_run(qpt, oracle, gpu):
    F, D = qpt.instances.random_instance(22, 48)
    betas = qpt.instances.geometric_ladder(8, 1.0, 100.0)
    inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=30, seed=9)
    with qpt.Pool([inst]) as pool:
        pool.run(10)
        pool.run(20)
        st = pool.state(0)
    ref = oracle.solve(F, D, betas, 30, seed=9)
    np.testing.assert_array_equal(st["perm"], ref["final_perm"])
    np.testing.assert_array_equal(st["E"], ref["final_E"])


def test_batch_matches_single_and_oracle(qpt, oracle, gpu):
    insts = []
    for k, n in enumerate([17, 64, 90, 131, 40]):
        F, D = qpt.instances.mixed_instance(100 + k, n, dense=(k % 2 == 0))
        betas = qpt.instances.geometric_ladder(4 + k, 1.0, 100.0)
        insts.append(qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=4 * (k + 1), seed=50 + k))
    batch = qpt.solve_batch(insts)
    for k, inst in enumerate(insts):
        single = qpt.solve(inst)
        ref = oracle.solve(inst.F, inst.D, inst.beta_ln, inst.sweeps, seed=inst.seed)
        assert batch[k]["best_cost"] == single["best_cost"] == ref["best_cost"]
        np.testing.assert_array_equal(batch[k]["best_perm"], ref["best_perm"])
        assert batch[k]["accepted"] == ref["accepted"]


def test_batch_mixes_one_shot_and_sweep_instances(qpt, oracle, gpu):
    """One-shot and p-dit instances share a pool (separate kernels in the same launch
    sequence); each ends where its own oracle run does."""
    insts = []
    for k, (n, one_shot) in enumerate([(40, True), (40, False), (100, True), (23, True),
                                       (180, True)]):
        F, D = qpt.instances.mixed_instance(120 + k, n, dense=(k % 2 == 0))
        betas = qpt.instances.geometric_ladder(3 + k, 1.0, 100.0) / max(1.0, n / 16.0)
        insts.append(qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=3 * (k + 1), seed=70 + k,
                                  one_shot=one_shot))
    batch = qpt.solve_batch(insts)
    for k, inst in enumerate(insts):
        ref = oracle.solve(inst.F, inst.D, inst.beta_ln, inst.sweeps, seed=inst.seed,
                           one_shot=inst.one_shot)
        print(f"instance {k}: n={inst.n} one_shot={inst.one_shot}", flush=True)
        assert batch[k]["best_cost"] == ref["best_cost"]
        np.testing.assert_array_equal(batch[k]["best_perm"], ref["best_perm"])
        assert batch[k]["accepted"] == ref["accepted"]
        assert batch[k]["attempts"] == ref["attempts"]


def test_timed_mode_same_trajectory(qpt, oracle, gpu):
    """Timed runs (events around every sweep launch) follow the same trajectory: one
    persistent launch for the 8 intervals of this latency-bound LDS pool, or 8 launches
    with persistent launches switched off."""
    F, D = qpt.instances.flat_c1()
    betas = qpt.instances.geometric_ladder(16, 1.0, 100.0)
    ref = oracle.solve(F, D, betas, 8, seed=4)
    for persist, want in (("1", 1), ("0", 8)):
        os.environ["QAP_PT_PERSIST"] = persist
        try:
            inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=8, seed=4)
            with qpt.Pool([inst]) as pool:
                ms, launches = pool.run(-1, timed=True)
                st = pool.state(0)
        finally:
            os.environ.pop("QAP_PT_PERSIST", None)
        assert launches == want and ms > 0, (persist, launches)
        np.testing.assert_array_equal(st["perm"], ref["final_perm"])
        np.testing.assert_array_equal(st["E"], ref["final_E"])


PERSIST_CASES = {
    # latency-bound LDS pools: intervals inside one launch vs one launch per interval
    "flat_c1_R16": (lambda q: q.instances.flat_c1(), 16, 60, {}),
    "grey_R8_T2_spe2_lds16": (lambda q: q.instances.grey_c2(), 8, 6, {"T": 2, "spe": 2}),
    "asym_n45_lds16_kick": (lambda q: q.instances.random_instance(23, 45, symmetric=False,
                                                                  zero_diag=False), 6, 40,
                            {"kick_after": 2, "env": {"QAP_PT_IMODE": "1"}}),
    "long_ladder_R70_T3": (lambda q: q.instances.random_instance(33, 16), 70, 12,
                           {"T": 3, "kick_after": 2}),
    "n100_256x1_T12": (lambda q: q.instances.mixed_instance(5, 100, True), 4, 10,
                       {"T": 12, "env": {"QAP_PT_SHAPE": "256,1"}}),
    "spec4_n40_T5": (lambda q: q.instances.random_instance(84, 40), 6, 30,
                     {"T": 5, "env": {"QAP_PT_SPEC": "4"}}),
    "spec8_n40_T5": (lambda q: q.instances.random_instance(87, 40), 6, 30,
                     {"T": 5, "env": {"QAP_PT_SPEC": "8"}}),
    "spec16_n40_T5": (lambda q: q.instances.random_instance(90, 40), 6, 30, {"T": 5}),
    # costs beyond 31 bits by the host's bound (max|F| sum|D| ~ 6e9) while the force matrix
    # still fits int32 LDS: the library keeps the counter protocol by itself
    "wide_costs_n48_R16": (lambda q: _wide_costs(93, 48), 16, 30, {}),
    # one ladder (the mirrored step's parallel bubble pass), three sweeps per exchange
    "spec16_n40_R12_spe3_kick": (lambda q: q.instances.random_instance(92, 40), 12, 36,
                                 {"spe": 3, "kick_after": 2}),
}


@pytest.mark.parametrize("name", list(PERSIST_CASES))
def test_persistent_launch_equals_interval_launches(qpt, oracle, gpu, name):
    """The exchange step inside a persistent launch -- mirrored by every workgroup from the
    replicas' tagged costs (ladders of up to 64 replicas), or run by the last replica of the
    instance to arrive (QAP_PT_XMIRROR=0, and every larger ladder set) -- gives the same
    trajectory as exchange launches between sweep launches, and all equal the oracle's; a
    goal stops each at the same interval."""
    build, R, sweeps, kw = PERSIST_CASES[name]
    F, D = build(qpt)
    n = F.shape[0]
    betas = qpt.instances.geometric_ladder(R, 1.0, 100.0) / max(1.0, n / 16.0)
    print(f"persist case {name}: n={n} R={R} sweeps={sweeps} kw={kw}", flush=True)
    env = dict(kw.pop("env", {}))
    modes = {"mirror": {"QAP_PT_PERSIST": "1"},
             "counter": {"QAP_PT_PERSIST": "1", "QAP_PT_XMIRROR": "0"},
             "launches": {"QAP_PT_PERSIST": "0"}}
    for mode, menv in modes.items():
        res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, sweeps,
                                      env=dict(env, **menv), **kw)
        _assert_same(res, st, ver, ref)
    free = oracle.solve(F, D, betas, sweeps, seed=7, **kw)
    goal = int(free["best_cost"] * 1.01)
    for mode, menv in modes.items():
        res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, sweeps, goal=goal,
                                      env=dict(env, **menv), **kw)
        assert res["sweeps_done"] == ref["megas_done"] * kw.get("spe", 1), mode
        _assert_same(res, st, ver, ref)


@pytest.mark.parametrize("xmirror", ["1", "0"])
def test_persistent_launch_shared_by_instances_one_stopping(qpt, oracle, gpu, xmirror):
    """Three instances of different n in one persistent launch (one kernel bucket), one of
    them with a goal it reaches partway: it leaves the launch at that exchange step while
    the others run on, each exchanging through its own replicas' words.  Every instance
    equals its own oracle run."""
    env = {"QAP_PT_PERSIST": "1", "QAP_PT_XMIRROR": xmirror}
    sweeps, R = 40, 8
    specs = [(94, 40), (95, 52), (96, 64)]
    data = [qpt.instances.random_instance(sd, n) for sd, n in specs]
    ladders = [qpt.instances.geometric_ladder(R, 1.0, 100.0) / max(1.0, n / 16.0)
               for _, n in specs]
    free = oracle.solve(*data[1], ladders[1], sweeps, seed=31)
    goals = [None, int(free["best_cost"] * 1.02), None]
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        insts = [qpt.Instance(F=F, D=D, beta_ln=b, sweeps=sweeps, seed=30 + k, goal=gl)
                 for k, ((F, D), b, gl) in enumerate(zip(data, ladders, goals))]
        with qpt.Pool(insts) as pool:
            shapes = {pool.shape(k) for k in range(3)}
            pool.run(-1)
            got = [(pool.result(k), pool.state(k)) for k in range(3)]
            ver = pool.verify()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    assert len(shapes) == 1, shapes  # one bucket: one launch holds all three
    for k, ((F, D), b, gl) in enumerate(zip(data, ladders, goals)):
        ref = oracle.solve(F, D, b, sweeps, seed=30 + k, goal=gl)
        res, st = got[k]
        _assert_same(res, st, ver, ref)
        assert res["sweeps_done"] == ref["megas_done"], k
    assert got[1][0]["sweeps_done"] < sweeps  # the goal ended instance 1 early


@pytest.mark.parametrize("xmirror", ["1", "0"])
def test_persistent_runs_split_across_launches(qpt, oracle, gpu, xmirror):
    """A run advanced in uneven pieces (persistent launches of 1, 2, 7 and the remaining
    intervals, each after an exchange launch) carries the ladder state from launch to
    launch: the mirrored step's copy is written back at the end of each launch as the
    exchange kernel would have left it.  Two ladders, kicks after two quiet steps."""
    F, D = qpt.instances.random_instance(89, 30)
    R, T, sweeps = 6, 2, 40
    betas = qpt.instances.geometric_ladder(R, 1.0, 100.0) / 2.0
    env = {"QAP_PT_PERSIST": "1", "QAP_PT_XMIRROR": xmirror}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=sweeps, T=T, seed=11, kick_after=2)
        with qpt.Pool([inst]) as pool:
            for piece in (1, 2, 7, 5):
                pool.run(piece)
            pool.run(-1)
            res, st = pool.result(0), pool.state(0)
            ver = pool.verify()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    ref = oracle.solve(F, D, betas, sweeps, T=T, seed=11, kick_after=2)
    assert ref["kicks"] > 0
    _assert_same(res, st, ver, ref)


def test_frozen_ladder_all_weights_zero(qpt, oracle, gpu):
    """beta so large that 20/beta < 1: at a local minimum (max dE = 0) every weight
    truncates to zero and the update must stay (the oracle's definition)."""
    F, D = qpt.instances.random_instance(31, 12)
    betas = np.array([5.0, 40.0, 400.0])
    res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, 30, seed=5)
    _assert_same(res, st, ver, ref)


def test_large_asym_n200_lds_shapes(qpt, oracle, gpu):
    F, D = qpt.instances.random_instance(32, 200, 0, 20, symmetric=False, zero_diag=False)
    betas = qpt.instances.geometric_ladder(3, 50.0, 5000.0)
    for shape in ("256,1", "128,2", "64,4"):
        res, st, ver, ref = _run_both(qpt, oracle, F, D, betas, 2, seed=6,
                                      env={"QAP_PT_SHAPE": shape})
        _assert_same(res, st, ver, ref)


def test_throughput_pool_more_replicas_than_cus(qpt, oracle, gpu):
    """Twelve instances of the 512-instance batch (n 30..150, dense and sparse flow, 32
    replicas each on T 100..10000): 384 replicas, more than the device's CUs, so the
    placements come from the tuned full-chip rates and several kernel buckets share each
    exchange interval.  Every instance's replica state equals its own oracle run."""
    insts = []
    for k, (F, D) in enumerate(qpt.instances.batch_c3()[:12]):
        insts.append(qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(32, 100.0, 10000.0),
                                  sweeps=6, seed=7000 + k))
    assert sum(x.R for x in insts) > 256  # MI355X: 256 CUs
    with qpt.Pool(insts) as pool:
        shapes = [pool.shape(k) for k in range(len(insts))]
        print("shapes:", shapes, flush=True)
        # every one of these sizes fits the int32 LDS placement, which a latency-bound
        # pool would take: a lazy placement shows the full-chip choice was made
        assert any(",lazy," in s for s in shapes), shapes
        pool.run(-1)
        got = [(pool.result(k), pool.state(k)) for k in range(len(insts))]
        fm, cm = pool.verify()
    for k, inst in enumerate(insts):
        ref = oracle.solve(inst.F, inst.D, inst.beta_ln, inst.sweeps, seed=inst.seed)
        print(f"instance {k}: n={inst.n} shape={shapes[k]}", flush=True)
        _assert_same(got[k][0], got[k][1], (fm, cm), ref)


def test_wide_instance_through_server_and_noalloc(qpt, oracle, gpu):
    """An instance beyond the int32 bounds through the serving loop (slot banks hold int32
    matrices, so it runs in a pool of its own beside banked requests) and through the
    allocation-free solve: both equal the allocating solve and the oracle."""
    F, D = _wide64(90, 48, symmetric=False, zero_diag=False)
    betas = qpt.instances.geometric_ladder(4, 1.0, 100.0) / 3.0 * _wide_scale(48)
    wide = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=6, seed=31)
    Fs, Ds = qpt.instances.mixed_instance(91, 40, dense=True)
    small = qpt.Instance(F=Fs, D=Ds, beta_ln=qpt.instances.geometric_ladder(4, 1.0, 100.0) / 2.5,
                         sweeps=6, seed=32)
    want = [qpt.solve(wide), qpt.solve(small)]
    ref = oracle.solve(F, D, betas, 6, seed=31)
    assert want[0]["best_cost"] == ref["best_cost"]
    np.testing.assert_array_equal(want[0]["best_perm"], ref["best_perm"])
    with qpt.Server(max_pools=4, max_batch=4) as srv:
        tickets = [srv.submit(wide), srv.submit(small)]
        got = [srv.wait(t, timeout_s=60.0) for t in tickets]
    for g, w in zip(got, want):
        assert g["status"] == qpt.QAP_SRV_DONE
        assert g["best_cost"] == w["best_cost"] and g["accepted"] == w["accepted"]
        np.testing.assert_array_equal(g["best_perm"], w["best_perm"])
    na = qpt.solve_noalloc(wide)
    assert na["best_cost"] == want[0]["best_cost"] and na["accepted"] == want[0]["accepted"]
    np.testing.assert_array_equal(na["best_perm"], want[0]["best_perm"])
