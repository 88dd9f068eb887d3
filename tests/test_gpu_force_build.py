"""Set-up force matrices on the matrix cores (force_mfma_kernel): every entry equals the
vector-ALU rebuild that qap_pt_pool_verify runs, right after pool creation, for entries
needing 1 to 4 signed base-256 digits (including the digit boundaries 127/128,
32639/32640, 8355711/8355712 and negative entries), every starting-permutation grouping
(a ladder's replicas, per-replica starts, a common init_perm, several ladders) and sizes
off the 64-wide tile grid.  Trajectory parity with the oracle (test_gpu_parity) covers the
same kernel on every case it runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inst(seed, n, mD, mF, signed):
    g = np.random.default_rng(seed)
    lo_d, lo_f = (-mD, -mF) if signed else (0, 0)
    D = g.integers(lo_d, mD + 1, (n, n)).astype(np.int32)
    F = g.integers(lo_f, mF + 1, (n, n)).astype(np.int32)
    # the boundary values themselves
    D[0, min(1, n - 1)] = mD
    F[min(1, n - 1), 0] = mF
    if signed:
        D[n - 1, 0] = -mD
        F[0, n - 1] = -mF
    return F, D


# (n, max|D|, max|F|, signed): digits of D x digits of F
MAGS = [
    (37, 9, 9, False),          # 1 x 1
    (64, 127, 100, True),       # 1 x 1 at the digit limit
    (65, 128, 20, True),        # 2 x 1
    (100, 32639, 3, True),      # 2 x 1 at the limit
    (70, 32640, 2, False),      # 3 x 1
    (24, 200, 30000, True),     # 2 x 2
    (12, 8355711, 1, True),     # 3 x 1 at the limit
    (10, 8355712, 1, False),    # 4 x 1
    (24, 1, 2_000_000, True),   # 1 x 3
    (129, 5, 40, False),        # 1 x 1, three tiles per side
]


@pytest.mark.parametrize("case", range(len(MAGS)))
def test_mfma_force_build_equals_rebuild(qpt, gpu, case):
    n, mD, mF, signed = MAGS[case]
    F, D = _inst(100 + case, n, mD, mF, signed)
    betas = np.geomspace(1e-9, 1e-7, 4)
    print(f"force build case {case}: n={n} max|D|={mD} max|F|={mF} signed={signed}", flush=True)
    for kw in ({}, {"T": 3}, {"per_replica_init": True, "T": 2},
               {"init_perm": np.random.default_rng(case).permutation(n).astype(np.int32), "T": 2}):
        inst = qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=0, **kw)
        with qpt.Pool([inst]) as pool:
            assert "hbm64" not in pool.shape(0), pool.shape(0)  # int32 force matrices
            fm, cm = pool.verify()
        assert (fm, cm) == (0, 0), (kw, fm, cm)


def test_mfma_force_build_batch_of_sizes(qpt, gpu):
    """Instances of uneven n in one pool (one build launch each) and the serving slots'
    load path."""
    insts = []
    for k, n in enumerate((1, 2, 15, 63, 64, 65, 127, 150)):
        F, D = _inst(200 + k, n, 50, 60, True)
        insts.append(qpt.Instance(F=F, D=D, beta_ln=np.geomspace(1e-6, 1e-5, 3), sweeps=0,
                                  T=2))
    with qpt.Pool(insts) as pool:
        assert pool.verify() == (0, 0)


def test_force_build_paths_give_the_same_run(qpt, gpu, monkeypatch):
    """The matrix-core build with costs read off the force matrices and the vector-ALU
    build with O(n^2) cost sums (QAP_PT_FORCE_BUILD=valu) start identical chains."""
    F, D = qpt.instances.random_instance(300, 90)
    inst = qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(6, 1.0, 100.0) / 6.0,
                        sweeps=5, T=2, seed=4)
    out = {}
    for mode in ("valu", "mfma"):
        monkeypatch.setenv("QAP_PT_FORCE_BUILD", mode)
        with qpt.Pool([inst]) as pool:
            E0 = pool.state(0)["E"].copy()
            pool.run(-1)
            out[mode] = (E0, pool.result(0), pool.verify())
    assert np.array_equal(out["valu"][0], out["mfma"][0])
    for k in ("best_cost", "accepted", "exchanges", "update_cells"):
        assert out["valu"][1][k] == out["mfma"][1][k], k
    assert out["mfma"][2] == (0, 0)
