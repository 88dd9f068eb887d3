"""Python binding of the MI355X parallel-tempering QAP annealer (libqap_pt.so).

The product is the C-ABI in ``include/qap_pt.h`` implemented by HIP kernels for
gfx950; this module is a thin ``ctypes`` layer over it for tests, benchmarks and
Python callers.  It mirrors the reference solve path (CPU-PIM/CPU_PIM.cpp:478-685):
flow/distance matrices, an explicit temperature ladder, sweeps and a seed go in; the
best permutation and its exact integer cost come out.  There is no CPU fallback:
loading fails loudly when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import instances  # noqa: F401  (re-exported helpers)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libqap_pt.so")

QAP_PT_OK = 0
QAP_PT_E_INVAL = -1
QAP_PT_E_NOMEM = -2
QAP_PT_E_DEVICE = -3
QAP_PT_E_OVERFLOW = -4
QAP_PT_E_BUSY = -5
QAP_PT_INIT_PER_REPLICA = 1
QAP_PT_ONE_SHOT = 2
NOALLOC_SCRATCH_BYTES = 64 << 20  # QAP_PT_NOALLOC_SCRATCH_BYTES
MAX_N = 1024
AUTO_MIN_EXP = -30
AUTO_STEPS = 32

# Every symbol the C header declares (checked by the CPU test suite).
EXPORTS = (
    "qap_pt_last_error",
    "qap_pt_device_count",
    "qap_pt_solve_hip",
    "qap_pt_solve_cpu",
    "qap_pt_solve_hip_noalloc",
    "qap_pt_solve_batch_hip",
    "qap_pt_pool_create",
    "qap_pt_pool_run",
    "qap_pt_pool_result",
    "qap_pt_pool_best_costs",
    "qap_pt_pool_state",
    "qap_pt_pool_verify",
    "qap_pt_pool_destroy",
    "qap_pt_kernel_shape",
    "qap_pt_pool_bucket",
    "qap_pt_pool_shape",
    "qap_pt_auto_beta",
    "qap_pt_auto_select",
    "qap_srv_create",
    "qap_srv_submit",
    "qap_srv_poll",
    "qap_srv_wait",
    "qap_srv_stats",
    "qap_srv_destroy",
)

QAP_SRV_QUEUED = 0
QAP_SRV_RUNNING = 1
QAP_SRV_DONE = 2

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U64P = C.POINTER(C.c_uint64)
_F64P = C.POINTER(C.c_double)


class QapProblem(C.Structure):
    _fields_ = [("n", C.c_int32), ("F", _I32P), ("D", _I32P)]


class QapPtParams(C.Structure):
    _fields_ = [
        ("R", C.c_int32),
        ("T", C.c_int32),
        ("beta_ln", _F64P),
        ("sweeps", C.c_int64),
        ("sweeps_per_exchange", C.c_int32),
        ("seed", C.c_uint64),
        ("kick_after", C.c_int32),
        ("has_goal", C.c_int32),
        ("goal", C.c_int64),
        ("timeout_s", C.c_double),
        ("flags", C.c_uint32),
        ("init_perm", _I32P),
    ]


class QapPtResult(C.Structure):
    _fields_ = [
        ("best_cost", C.c_int64),
        ("best_perm", _I32P),
        ("attempts", C.c_uint64),
        ("accepted", C.c_uint64),
        ("exchanges", C.c_uint64),
        ("kicks", C.c_uint64),
        ("update_cells", C.c_uint64),
        ("flush_cells", C.c_uint64),
        ("sweeps_done", C.c_int64),
        ("seconds", C.c_double),
    ]


class QapSrvConfig(C.Structure):
    _fields_ = [
        ("max_pools", C.c_int32),
        ("max_batch", C.c_int32),
        ("device_budget", C.c_int64),
        ("reserve_bytes", C.c_int64),
        ("batch_window_s", C.c_double),
    ]


class QapError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"qap_pt error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libqap_pt.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    lib.qap_pt_last_error.restype = C.c_char_p
    lib.qap_pt_device_count.restype = C.c_int
    lib.qap_pt_solve_hip.argtypes = [C.POINTER(QapProblem), C.POINTER(QapPtParams),
                                     C.POINTER(QapPtResult)]
    lib.qap_pt_solve_cpu.argtypes = [C.POINTER(QapProblem), C.POINTER(QapPtParams),
                                     C.POINTER(QapPtResult)]
    lib.qap_pt_solve_hip_noalloc.argtypes = [C.POINTER(QapProblem), C.POINTER(QapPtParams),
                                             C.POINTER(QapPtResult), C.c_void_p]
    lib.qap_pt_solve_batch_hip.argtypes = [C.c_int32, C.POINTER(QapProblem),
                                           C.POINTER(QapPtParams), C.POINTER(QapPtResult)]
    lib.qap_pt_pool_create.argtypes = [C.c_int32, C.POINTER(QapProblem),
                                       C.POINTER(QapPtParams), C.POINTER(C.c_void_p)]
    lib.qap_pt_pool_run.argtypes = [C.c_void_p, C.c_int64, _F64P, _I64P]
    lib.qap_pt_pool_result.argtypes = [C.c_void_p, C.c_int32, C.POINTER(QapPtResult)]
    lib.qap_pt_pool_state.argtypes = [C.c_void_p, C.c_int32, _I32P, _I64P, _I64P, _I32P,
                                      _U64P, _I32P]
    lib.qap_pt_pool_best_costs.argtypes = [C.c_void_p, _I64P]
    lib.qap_pt_pool_verify.argtypes = [C.c_void_p, _I64P, _I64P]
    lib.qap_pt_pool_destroy.argtypes = [C.c_void_p]
    lib.qap_pt_kernel_shape.argtypes = [C.c_int32, C.c_int64, C.c_char_p, C.c_int32]
    lib.qap_pt_pool_shape.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]
    lib.qap_pt_pool_bucket.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, _I64P,
                                       _F64P, _I64P]
    lib.qap_pt_auto_beta.argtypes = [C.POINTER(QapProblem), C.POINTER(QapPtParams), C.c_double,
                                     C.c_double, _F64P, _F64P, _F64P]
    lib.qap_pt_auto_select.argtypes = [_F64P, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                       _F64P, _F64P]
    lib.qap_srv_create.argtypes = [C.POINTER(QapSrvConfig), C.POINTER(C.c_void_p)]
    lib.qap_srv_submit.argtypes = [C.c_void_p, C.POINTER(QapProblem), C.POINTER(QapPtParams),
                                   C.POINTER(C.c_uint64)]
    lib.qap_srv_poll.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(QapPtResult),
                                 C.POINTER(C.c_int32)]
    lib.qap_srv_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(QapPtResult),
                                 C.POINTER(C.c_int32)]
    lib.qap_srv_stats.argtypes = [C.c_void_p, _I64P, _I64P, _I64P, _I64P]
    lib.qap_srv_destroy.argtypes = [C.c_void_p]
    _lib = lib
    return lib


def _check(rc: int):
    if rc != QAP_PT_OK:
        raise QapError(rc, load_library().qap_pt_last_error().decode())


def device_count() -> int:
    return int(load_library().qap_pt_device_count())


def kernel_shape(n: int, replicas: int = 1) -> str:
    """Shape and placement for a symmetric size-n instance in a pool of `replicas`."""
    buf = C.create_string_buffer(64)
    _check(load_library().qap_pt_kernel_shape(n, replicas, buf, 64))
    return buf.value.decode()


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class Instance:
    """One QAP instance plus its annealing parameters (the reference's flags)."""
    F: np.ndarray
    D: np.ndarray
    beta_ln: np.ndarray
    sweeps: int
    R: int = 0
    T: int = 1
    sweeps_per_exchange: int = 1
    seed: int = 0
    kick_after: int = 100
    goal: int | None = None
    timeout_s: float = 0.0
    per_replica_init: bool = False
    init_perm: np.ndarray | None = None
    one_shot: bool = False  # CUPIM's one-shot parallel-proposal steps (a different chain)
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.F = np.ascontiguousarray(self.F, dtype=np.int32)
        self.D = np.ascontiguousarray(self.D, dtype=np.int32)
        self.beta_ln = np.ascontiguousarray(self.beta_ln, dtype=np.float64)
        if self.R == 0:
            self.R = int(self.beta_ln.size)
        if self.init_perm is not None:
            self.init_perm = np.ascontiguousarray(self.init_perm, dtype=np.int32)

    @property
    def n(self) -> int:
        return int(self.F.shape[0])

    def c_structs(self):
        prob = QapProblem(self.n, _ptr(self.F, C.c_int32), _ptr(self.D, C.c_int32))
        par = QapPtParams()
        par.R, par.T = self.R, self.T
        par.beta_ln = _ptr(self.beta_ln, C.c_double)
        par.sweeps = int(self.sweeps)
        par.sweeps_per_exchange = int(self.sweeps_per_exchange)
        par.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        par.kick_after = int(self.kick_after)
        par.has_goal = 0 if self.goal is None else 1
        par.goal = 0 if self.goal is None else int(self.goal)
        par.timeout_s = float(self.timeout_s)
        par.flags = (QAP_PT_INIT_PER_REPLICA if self.per_replica_init else 0) | \
            (QAP_PT_ONE_SHOT if self.one_shot else 0)
        par.init_perm = (_ptr(self.init_perm, C.c_int32) if self.init_perm is not None
                         else C.cast(None, _I32P))
        return prob, par


def auto_select(avg_energy, R: int, lower_prop: float = 0.1,
                higher_prop: float = 0.9) -> tuple[float, float]:
    """qap_pt_auto_select: the reference's ladder rule (CPU_PIM.cpp:431-473) over mean
    final costs at exp2-unit betas 2^(AUTO_MIN_EXP + i).  Returns (beta0_ln, ratio)."""
    avg = np.ascontiguousarray(avg_energy, dtype=np.float64)
    b0, m = C.c_double(), C.c_double()
    _check(load_library().qap_pt_auto_select(_ptr(avg, C.c_double), avg.size, lower_prop,
                                             higher_prop, R, C.byref(b0), C.byref(m)))
    return b0.value, m.value


def auto_beta(F, D, R: int, T: int = 1, sweeps_per_exchange: int = 1, seed: int = 0,
              lower_prop: float = 0.1, higher_prop: float = 0.9, per_replica_init=False,
              one_shot=False):
    """qap_pt_auto_beta: the automatic ladder scan (CPU_PIM.cpp:366-476) as one batched
    run (one_shot: the trials run one-shot steps).  Returns (beta0_ln, ratio,
    avg_energy[AUTO_STEPS]); the ladder is then beta_ln[l] = beta0_ln / ratio**(R - 1 - l)."""
    inst = Instance(F=F, D=D, beta_ln=np.ones(R), sweeps=0, T=T,
                    sweeps_per_exchange=sweeps_per_exchange, seed=seed,
                    per_replica_init=per_replica_init, one_shot=one_shot)
    prob, par = inst.c_structs()
    avg = np.zeros(AUTO_STEPS, dtype=np.float64)
    b0, m = C.c_double(), C.c_double()
    _check(load_library().qap_pt_auto_beta(C.byref(prob), C.byref(par), lower_prop, higher_prop,
                                           _ptr(avg, C.c_double), C.byref(b0), C.byref(m)))
    return b0.value, m.value, avg


def ladder_from(beta0_ln: float, ratio: float, R: int) -> np.ndarray:
    """The reference's ladder (CPU_PIM.cpp:113-118): replica R-1 at beta0, each colder
    rung divided by ratio in turn; returned hottest first."""
    return instances.beta_ladder_from_flags(R, beta0_ln, ratio)


def _result_dict(res: QapPtResult, perm: np.ndarray) -> dict:
    return {
        "best_cost": int(res.best_cost),
        "best_perm": perm,
        "attempts": int(res.attempts),
        "accepted": int(res.accepted),
        "exchanges": int(res.exchanges),
        "kicks": int(res.kicks),
        "update_cells": int(res.update_cells),
        "flush_cells": int(res.flush_cells),
        "sweeps_done": int(res.sweeps_done),
        "seconds": float(res.seconds),
    }


def solve(inst: Instance) -> dict:
    """qap_pt_solve_hip: one instance, run to inst.sweeps."""
    lib = load_library()
    prob, par = inst.c_structs()
    perm = np.zeros(inst.n, dtype=np.int32)
    res = QapPtResult()
    res.best_perm = _ptr(perm, C.c_int32)
    _check(lib.qap_pt_solve_hip(C.byref(prob), C.byref(par), C.byref(res)))
    return _result_dict(res, perm)


def solve_cpu(inst: Instance, threads: int | None = None) -> dict:
    """qap_pt_solve_cpu: the same solve on the host's cores (no device call); identical
    results to `solve`.  threads sets QAP_PT_CPU_THREADS for the call (None = the
    environment's value, else every core the process may run on)."""
    lib = load_library()
    prob, par = inst.c_structs()
    perm = np.zeros(inst.n, dtype=np.int32)
    res = QapPtResult()
    res.best_perm = _ptr(perm, C.c_int32)
    old = os.environ.get("QAP_PT_CPU_THREADS")
    if threads is not None:
        os.environ["QAP_PT_CPU_THREADS"] = str(int(threads))
    try:
        _check(lib.qap_pt_solve_cpu(C.byref(prob), C.byref(par), C.byref(res)))
    finally:
        if threads is not None:
            if old is None:
                os.environ.pop("QAP_PT_CPU_THREADS", None)
            else:
                os.environ["QAP_PT_CPU_THREADS"] = old
    return _result_dict(res, perm)


def solve_noalloc(inst: Instance, stream: int | None = None) -> dict:
    """qap_pt_solve_hip_noalloc: the solve of `solve`, with every device buffer in the
    library's fixed per-GPU region (no device allocation); kernels on `stream` (a
    hipStream_t as an integer, None = the null stream)."""
    lib = load_library()
    prob, par = inst.c_structs()
    perm = np.zeros(inst.n, dtype=np.int32)
    res = QapPtResult()
    res.best_perm = _ptr(perm, C.c_int32)
    _check(lib.qap_pt_solve_hip_noalloc(C.byref(prob), C.byref(par), C.byref(res),
                                        C.c_void_p(stream)))
    return _result_dict(res, perm)


def solve_batch(insts: list[Instance]) -> list[dict]:
    """qap_pt_solve_batch_hip: every instance in the same device launches."""
    lib = load_library()
    k = len(insts)
    probs = (QapProblem * k)()
    pars = (QapPtParams * k)()
    ress = (QapPtResult * k)()
    perms = []
    for x, inst in enumerate(insts):
        probs[x], pars[x] = inst.c_structs()
        p = np.zeros(inst.n, dtype=np.int32)
        perms.append(p)
        ress[x].best_perm = _ptr(p, C.c_int32)
    _check(lib.qap_pt_solve_batch_hip(k, probs, pars, ress))
    return [_result_dict(ress[x], perms[x]) for x in range(k)]


class Pool:
    """Staged pool (the reference's qap::ReplicaPool, qap.cu:28-303)."""

    def __init__(self, insts: list[Instance]):
        self.lib = load_library()
        self.insts = list(insts)
        k = len(self.insts)
        self._probs = (QapProblem * k)()
        self._pars = (QapPtParams * k)()
        for x, inst in enumerate(self.insts):
            self._probs[x], self._pars[x] = inst.c_structs()
        self.handle = C.c_void_p()
        _check(self.lib.qap_pt_pool_create(k, self._probs, self._pars, C.byref(self.handle)))

    def run(self, sweeps: int = -1, timed: bool = False):
        """Advance every instance; with timed=True returns (kernel_ms, launches)."""
        if timed:
            ms = C.c_double(0.0)
            nl = C.c_int64(0)
            _check(self.lib.qap_pt_pool_run(self.handle, int(sweeps), C.byref(ms), C.byref(nl)))
            return float(ms.value), int(nl.value)
        _check(self.lib.qap_pt_pool_run(self.handle, int(sweeps), None, None))
        return None

    def result(self, k: int = 0) -> dict:
        perm = np.zeros(self.insts[k].n, dtype=np.int32)
        res = QapPtResult()
        res.best_perm = _ptr(perm, C.c_int32)
        _check(self.lib.qap_pt_pool_result(self.handle, k, C.byref(res)))
        return _result_dict(res, perm)

    def best_costs(self) -> np.ndarray:
        """Every instance's best cost so far (one device copy)."""
        out = np.zeros(len(self.insts), dtype=np.int64)
        _check(self.lib.qap_pt_pool_best_costs(self.handle, _ptr(out, C.c_int64)))
        return out

    def state(self, k: int = 0) -> dict:
        inst = self.insts[k]
        rt, n = inst.R * inst.T, inst.n
        perm = np.zeros((rt, n), dtype=np.int32)
        bperm = np.zeros((rt, n), dtype=np.int32)
        E = np.zeros(rt, dtype=np.int64)
        bE = np.zeros(rt, dtype=np.int64)
        pos = np.zeros(rt, dtype=np.int32)
        acc = np.zeros(rt, dtype=np.uint64)
        _check(self.lib.qap_pt_pool_state(self.handle, k, _ptr(perm, C.c_int32),
                                          _ptr(E, C.c_int64), _ptr(bE, C.c_int64),
                                          _ptr(pos, C.c_int32), _ptr(acc, C.c_uint64),
                                          _ptr(bperm, C.c_int32)))
        return {"perm": perm, "E": E, "bestE": bE, "pos": pos, "accepted": acc,
                "best_perms": bperm}

    def shape(self, k: int = 0) -> str:
        buf = C.create_string_buffer(96)
        _check(self.lib.qap_pt_pool_shape(self.handle, k, buf, 96))
        return buf.value.decode()

    def buckets(self) -> list[dict]:
        """Kernel buckets: shape, replicas, and device time / launches of timed runs."""
        out, b = [], 0
        buf = C.create_string_buffer(96)
        reps, ms, nl = C.c_int64(0), C.c_double(0.0), C.c_int64(0)
        while self.lib.qap_pt_pool_bucket(self.handle, b, buf, 96, C.byref(reps), C.byref(ms),
                                          C.byref(nl)) == QAP_PT_OK:
            out.append({"shape": buf.value.decode(), "replicas": int(reps.value),
                        "kernel_ms": float(ms.value), "launches": int(nl.value)})
            b += 1
        return out

    def verify(self) -> tuple[int, int]:
        fm = C.c_int64(0)
        cm = C.c_int64(0)
        _check(self.lib.qap_pt_pool_verify(self.handle, C.byref(fm), C.byref(cm)))
        return int(fm.value), int(cm.value)

    def close(self):
        if self.handle:
            self.lib.qap_pt_pool_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop (qap_srv_*): continuous batching of solve requests on one GPU."""

    def __init__(self, max_pools=4, max_batch=64, device_budget=0, reserve_bytes=1 << 30,
                 batch_window_s=0.0):
        self.lib = load_library()
        cfg = QapSrvConfig(int(max_pools), int(max_batch), int(device_budget),
                           int(reserve_bytes), float(batch_window_s))
        self.handle = C.c_void_p()
        _check(self.lib.qap_srv_create(C.byref(cfg), C.byref(self.handle)))
        self._n = {}

    def submit(self, inst: Instance) -> int:
        prob, par = inst.c_structs()
        t = C.c_uint64(0)
        _check(self.lib.qap_srv_submit(self.handle, C.byref(prob), C.byref(par), C.byref(t)))
        self._n[int(t.value)] = inst.n
        return int(t.value)

    def _final(self, t, res, perm, st):
        self._n.pop(t, None)
        out = _result_dict(res, perm)
        out["status"] = int(st.value)
        return out

    def poll(self, t: int):
        """None while queued/running; the result dict (with 'status') once final."""
        perm = np.zeros(self._n[t], dtype=np.int32)
        res = QapPtResult()
        res.best_perm = _ptr(perm, C.c_int32)
        st = C.c_int32(0)
        _check(self.lib.qap_srv_poll(self.handle, t, C.byref(res), C.byref(st)))
        if st.value in (QAP_SRV_QUEUED, QAP_SRV_RUNNING):
            return None
        return self._final(t, res, perm, st)

    def wait(self, t: int, timeout_s: float = 0.0):
        perm = np.zeros(self._n[t], dtype=np.int32)
        res = QapPtResult()
        res.best_perm = _ptr(perm, C.c_int32)
        st = C.c_int32(0)
        _check(self.lib.qap_srv_wait(self.handle, t, float(timeout_s), C.byref(res),
                                     C.byref(st)))
        if st.value in (QAP_SRV_QUEUED, QAP_SRV_RUNNING):
            return None
        return self._final(t, res, perm, st)

    def stats(self) -> dict:
        q, r, d, b = (C.c_int64(0) for _ in range(4))
        _check(self.lib.qap_srv_stats(self.handle, C.byref(q), C.byref(r), C.byref(d),
                                      C.byref(b)))
        return {"queued": q.value, "running": r.value, "done": d.value, "device_bytes": b.value}

    def close(self):
        if self.handle:
            self.lib.qap_srv_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
