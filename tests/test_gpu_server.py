"""Serving loop on the GPU: requests admitted while others run, each result the same
alone and in the mix (bit for bit, same as the oracle), admission gated on memory."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(qpt, count, seed0=500):
    sizes = qpt.instances.served_sizes(count, seed=seed0)
    out = []
    for k, n in enumerate(sizes):
        n = int(min(n, 140))
        F, D = qpt.instances.mixed_instance(seed0 + k, n, dense=(k % 2 == 0))
        betas = qpt.instances.geometric_ladder(8, 1.0, 100.0) / max(1.0, n / 16.0)
        out.append(qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=6 + 2 * (k % 3), seed=900 + k))
    return out


def test_results_same_alone_and_in_the_mix(qpt, oracle, gpu):
    reqs = _requests(qpt, 10)
    alone = [qpt.solve(r) for r in reqs]
    with qpt.Server(max_pools=3, max_batch=4) as srv:
        tickets = []
        for k, r in enumerate(reqs):          # staggered arrivals: later ones join while
            tickets.append(srv.submit(r))     # earlier pools are on the device
            if k % 3 == 2:
                time.sleep(0.002)
        mixed = [srv.wait(t, timeout_s=60.0) for t in tickets]
    for k, r in enumerate(reqs):
        assert mixed[k] is not None and mixed[k]["status"] == qpt.QAP_SRV_DONE
        assert mixed[k]["best_cost"] == alone[k]["best_cost"]
        np.testing.assert_array_equal(mixed[k]["best_perm"], alone[k]["best_perm"])
        assert mixed[k]["accepted"] == alone[k]["accepted"]
        ref = oracle.solve(r.F, r.D, r.beta_ln, r.sweeps, seed=r.seed)
        assert mixed[k]["best_cost"] == ref["best_cost"]
        assert mixed[k]["seconds"] > 0


def test_budget_serialises_admission(qpt, oracle, gpu):
    """A budget that holds one request at a time still completes every request, in
    order, with the same results."""
    reqs = _requests(qpt, 4, seed0=700)
    from ctypes import byref  # noqa: F401
    lib = qpt.load_library()
    import ctypes as C
    sizes = []
    for r in reqs:
        prob, par = r.c_structs()
        sizes.append(r)
    del lib, C, sizes
    alone = [qpt.solve(r) for r in reqs]
    biggest = max(4 * r.R * r.T * r.n * ((r.n + 3) // 4 * 4) + 20 * r.n * r.n for r in reqs)
    with qpt.Server(max_pools=4, max_batch=4, device_budget=int(biggest * 1.3)) as srv:
        tickets = [srv.submit(r) for r in reqs]
        done = [srv.wait(t, timeout_s=60.0) for t in tickets]
        st = srv.stats()
    assert st["device_bytes"] == 0
    for k in range(len(reqs)):
        assert done[k]["status"] == qpt.QAP_SRV_DONE
        assert done[k]["best_cost"] == alone[k]["best_cost"]
