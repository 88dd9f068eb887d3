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


def test_banks_mixed_classes_goal_flight_and_empty_run(qpt, gpu):
    """Requests of several bank classes (n up to 300), more of them than a bank has slots,
    one with a goal (halts early), one with more replicas than any bank slot (runs as a
    pool of its own) and one with no sweeps: every result equals a standalone solve."""
    reqs = []
    for k, n in enumerate([20, 70, 130, 200, 300, 45, 64, 129, 90, 33]):
        F, D = qpt.instances.mixed_instance(800 + k, n, dense=(k % 2 == 1))
        betas = qpt.instances.geometric_ladder(8, 1.0, 100.0) / max(1.0, n / 16.0)
        reqs.append(qpt.Instance(F=F, D=D, beta_ln=betas, sweeps=4 + 2 * (k % 3), seed=1200 + k))
    F, D = qpt.instances.mixed_instance(850, 40, dense=True)
    free = qpt.solve(qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(8, 1.0, 20.0),
                                  sweeps=40, seed=1300))
    reqs.append(qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(8, 1.0, 20.0),
                             sweeps=40, seed=1300, goal=int(free["best_cost"] * 1.05)))
    reqs.append(qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(16, 1.0, 20.0),
                             sweeps=6, T=5, seed=1301))                       # 80 replicas
    reqs.append(qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(8, 1.0, 20.0),
                             sweeps=0, seed=1302))                            # no sweeps
    alone = [qpt.solve(r) for r in reqs]
    with qpt.Server(max_pools=8, max_batch=3) as srv:
        tickets = [srv.submit(r) for r in reqs]
        got = [srv.wait(t, timeout_s=120.0) for t in tickets]
    for k in range(len(reqs)):
        print(f"request {k}: n={reqs[k].n} status={got[k] and got[k]['status']}", flush=True)
        assert got[k] is not None and got[k]["status"] == qpt.QAP_SRV_DONE
        assert got[k]["best_cost"] == alone[k]["best_cost"]
        np.testing.assert_array_equal(got[k]["best_perm"], alone[k]["best_perm"])
        assert got[k]["accepted"] == alone[k]["accepted"]
        assert got[k]["sweeps_done"] == alone[k]["sweeps_done"]
        assert got[k]["exchanges"] == alone[k]["exchanges"]
    assert alone[-4]["sweeps_done"] < 40          # the goal stopped it early


def test_one_shot_requests_run_beside_sweep_requests(qpt, oracle, gpu):
    """One-shot requests (pools of their own) and p-dit requests (slot banks) in one stream;
    every result equals the oracle's run of its own dynamics."""
    reqs = _requests(qpt, 6, seed0=700)
    for k in (1, 3, 4):
        reqs[k].one_shot = True
    with qpt.Server(max_pools=4, max_batch=4) as srv:
        tickets = [srv.submit(r) for r in reqs]
        got = [srv.wait(t, timeout_s=60.0) for t in tickets]
    for k, r in enumerate(reqs):
        assert got[k] is not None and got[k]["status"] == qpt.QAP_SRV_DONE, k
        ref = oracle.solve(r.F, r.D, r.beta_ln, r.sweeps, seed=r.seed, one_shot=r.one_shot)
        assert got[k]["best_cost"] == ref["best_cost"], k
        np.testing.assert_array_equal(got[k]["best_perm"], ref["best_perm"])
        assert got[k]["accepted"] == ref["accepted"]


def test_submit_returns_while_large_requests_are_admitted(qpt, gpu):
    """Three n = 1024 requests (64 replicas each: host preparation, uploads, force-matrix
    builds of about 0.1 s each) are being admitted while small requests arrive.  Submit
    only copies the request and queues it under a short lock -- the worker does device
    work with the lock released -- so every submit returns within milliseconds, and the
    latencies count from the moment submit was called."""
    F, D = qpt.instances.large_c5(1024)
    big = [qpt.Instance(F=F, D=D, beta_ln=qpt.instances.geometric_ladder(64, 300.0, 30000.0),
                        sweeps=1, seed=40 + k) for k in range(3)]
    small = _requests(qpt, 12, seed0=950)
    alone = [qpt.solve(r) for r in small]
    with qpt.Server(max_pools=8, max_batch=4) as srv:
        tb = [srv.submit(r) for r in big]
        time.sleep(0.01)  # the worker is now preparing the first large request
        took, ts = [], []
        for r in small:
            t0 = time.perf_counter()
            ts.append(srv.submit(r))
            took.append(time.perf_counter() - t0)
        got = [srv.wait(t, timeout_s=120.0) for t in ts]
        gb = [srv.wait(t, timeout_s=120.0) for t in tb]
    print("submit seconds:", [round(x, 4) for x in took], flush=True)
    assert max(took) < 0.02, took
    for k in range(len(small)):
        assert got[k]["status"] == qpt.QAP_SRV_DONE
        assert got[k]["best_cost"] == alone[k]["best_cost"]
        assert got[k]["seconds"] > 0
    for g in gb:
        assert g["status"] == qpt.QAP_SRV_DONE
