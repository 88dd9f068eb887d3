"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker (or the timed CPU baseline).  See qap_oracle.h for
the reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "cpu_pim_ref")

MODE_EXACT = 0
MODE_REFFP = 1
FLAG_INIT_PER_REPLICA = 1
FLAG_ONE_SHOT = 2

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)
_F64P = C.POINTER(C.c_double)


class _Problem(C.Structure):
    _fields_ = [("n", C.c_int32), ("F", _I32P), ("D", _I32P)]


class _Params(C.Structure):
    _fields_ = [("R", C.c_int32), ("T", C.c_int32), ("beta_ln", _F64P), ("sweeps", C.c_int64),
                ("spe", C.c_int32), ("seed", C.c_uint64), ("kick_after", C.c_int32),
                ("has_goal", C.c_int32), ("goal", C.c_int64), ("flags", C.c_int32),
                ("init_perm", _I32P)]


class _Result(C.Structure):
    _fields_ = [("best_cost", C.c_int64), ("best_perm", _I32P), ("attempts", C.c_uint64),
                ("accepted", C.c_uint64), ("exchanges", C.c_uint64), ("kicks", C.c_uint64),
                ("megas_done", C.c_int64), ("final_perm", _I32P), ("final_E", _I64P),
                ("final_bestE", _I64P), ("final_pos", _I32P), ("final_accepted", _U64P),
                ("best_perms", _I32P), ("traj_E", _I64P), ("traj_cap", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_philox4x32_10.argtypes = [_U32P, _U32P, _U32P]
        L.oracle_weight.argtypes = [C.c_float, C.c_int32]
        L.oracle_weight.restype = C.c_uint32
        L.oracle_weight64.argtypes = [C.c_float, C.c_int64]
        L.oracle_weight64.restype = C.c_uint32
        L.oracle_force_matrix.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I32P]
        L.oracle_energy.argtypes = [C.c_int32, _I32P, _I32P, _I32P]
        L.oracle_energy.restype = C.c_int64
        L.oracle_delta.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P, C.c_int32,
                                   C.c_int32]
        L.oracle_delta.restype = C.c_int64
        L.oracle_force_update.argtypes = [C.c_int32, _I32P, _I32P, _I32P, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_int32]
        L.oracle_attempt_exact.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P, _I64P,
                                           C.c_int32, C.c_float, C.c_double, C.c_uint32, _I32P,
                                           _U32P]
        L.oracle_attempt_exact.restype = C.c_int32
        L.oracle_one_shot_step.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P,
                                           _I64P, C.c_float, C.c_double, C.c_uint64,
                                           C.c_uint32, C.c_uint32, _I32P,
                                           C.POINTER(C.c_uint64), _U32P, _I32P, _I32P]
        L.oracle_one_shot_step.restype = C.c_int32
        L.oracle_beta2.argtypes = [C.c_double]
        L.oracle_beta2.restype = C.c_double
        L.oracle_init_perm.argtypes = [C.c_int32, C.c_uint64, C.c_uint32, _I32P]
        L.oracle_solve.argtypes = [C.POINTER(_Problem), C.POINTER(_Params), C.POINTER(_Result),
                                   C.c_int32, C.c_int32]
        L.oracle_bounds.argtypes = [C.c_int32, _I32P, _I32P, _I64P, _I64P]
        L.oracle_exchange_step.argtypes = [C.c_int32, _I64P, _I64P, _I32P, _I32P, _I32P, _I64P,
                                           C.c_int32, _I32P]
        L.oracle_exchange_step.restype = C.c_int32
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return [int(x) for x in out]


def weight(beta2f: float, d: int) -> int:
    return int(lib().oracle_weight(C.c_float(beta2f), int(d)))


def weight64(beta2f: float, d: int) -> int:
    return int(lib().oracle_weight64(C.c_float(beta2f), int(d)))


def beta2(beta_ln: float) -> float:
    return float(lib().oracle_beta2(beta_ln))


def force_matrix(F, D, perm):
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    p = np.ascontiguousarray(perm, np.int32)
    n = F.shape[0]
    I = np.zeros((n, n), np.int32)
    lib().oracle_force_matrix(n, _p(F, C.c_int32), _p(D, C.c_int32), _p(p, C.c_int32),
                              _p(I, C.c_int32))
    return I


def energy(F, D, perm) -> int:
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    p = np.ascontiguousarray(perm, np.int32)
    return int(lib().oracle_energy(F.shape[0], _p(F, C.c_int32), _p(D, C.c_int32),
                                   _p(p, C.c_int32)))


def delta(F, D, perm, I, i, b) -> int:
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    p = np.ascontiguousarray(perm, np.int32)
    pinv = np.argsort(p).astype(np.int32)
    I = np.ascontiguousarray(I, np.int32)
    return int(lib().oracle_delta(F.shape[0], _p(F, C.c_int32), _p(D, C.c_int32),
                                  _p(p, C.c_int32), _p(pinv, C.c_int32), _p(I, C.c_int32), i, b))


def force_update(F, D, I, i, j, a, b):
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    I = np.ascontiguousarray(I, np.int32).copy()
    lib().oracle_force_update(F.shape[0], _p(F, C.c_int32), _p(D, C.c_int32), _p(I, C.c_int32),
                              i, j, a, b)
    return I


def attempt_exact(F, D, perm, I, E, i, beta2f, c20, u):
    """One exact p-dit update on copies; returns (sel, dE, w, perm', I', E')."""
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    n = F.shape[0]
    p = np.ascontiguousarray(perm, np.int32).copy()
    pinv = np.argsort(p).astype(np.int32)
    I = np.ascontiguousarray(I, np.int32).copy()
    e = C.c_int64(int(E))
    dE = np.zeros(n, np.int32)
    w = np.zeros(n, np.uint32)
    sel = lib().oracle_attempt_exact(n, _p(F, C.c_int32), _p(D, C.c_int32), _p(p, C.c_int32),
                                     _p(pinv, C.c_int32), _p(I, C.c_int32), C.byref(e), i,
                                     C.c_float(beta2f), C.c_double(c20), C.c_uint32(u),
                                     _p(dE, C.c_int32), _p(w, C.c_uint32))
    return int(sel), dE, w, p, I, int(e.value)


def one_shot_step(F, D, perm, I, E, beta2f, c20, seed, step, stream):
    """One one-shot parallel-proposal step on copies; returns
    (applied, proposals, won, perm', I', E')."""
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    n = F.shape[0]
    p = np.ascontiguousarray(perm, np.int32).copy()
    pinv = np.argsort(p).astype(np.int32)
    I = np.ascontiguousarray(I, np.int32).copy()
    e = C.c_int64(int(E))
    work = np.zeros(4 * n, np.int32)
    owner = np.zeros(n, np.uint64)
    w = np.zeros(n, np.uint32)
    tgt = np.zeros(n, np.int32)
    won = np.zeros(n, np.int32)
    m = lib().oracle_one_shot_step(n, _p(F, C.c_int32), _p(D, C.c_int32), _p(p, C.c_int32),
                                   _p(pinv, C.c_int32), _p(I, C.c_int32), C.byref(e),
                                   C.c_float(beta2f), C.c_double(c20), C.c_uint64(seed),
                                   C.c_uint32(step), C.c_uint32(stream), _p(work, C.c_int32),
                                   _p(owner, C.c_uint64), _p(w, C.c_uint32),
                                   _p(tgt, C.c_int32), _p(won, C.c_int32))
    return int(m), tgt, won.astype(bool), p, I, int(e.value)


def init_perm(n, seed, stream):
    p = np.zeros(n, np.int32)
    lib().oracle_init_perm(n, C.c_uint64(seed), C.c_uint32(stream), _p(p, C.c_int32))
    return p


def bounds(F, D):
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    fb = C.c_int64()
    db = C.c_int64()
    lib().oracle_bounds(F.shape[0], _p(F, C.c_int32), _p(D, C.c_int32), C.byref(fb),
                        C.byref(db))
    return int(fb.value), int(db.value)


def solve(F, D, beta_ln, sweeps, R=None, T=1, spe=1, seed=0, kick_after=100, goal=None,
          per_replica_init=False, init_perm=None, mode=MODE_EXACT, nthreads=1, traj=False,
          one_shot=False):
    """Full PT-SA solve; returns a dict with the result and the final replica state."""
    F = np.ascontiguousarray(F, np.int32)
    D = np.ascontiguousarray(D, np.int32)
    beta_ln = np.ascontiguousarray(beta_ln, np.float64)
    n = F.shape[0]
    R = int(beta_ln.size) if R is None else int(R)
    RT = R * T
    prob = _Problem(n, _p(F, C.c_int32), _p(D, C.c_int32))
    par = _Params()
    par.R, par.T = R, T
    par.beta_ln = _p(beta_ln, C.c_double)
    par.sweeps, par.spe = int(sweeps), int(spe)
    par.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    par.kick_after = int(kick_after)
    par.has_goal = 0 if goal is None else 1
    par.goal = 0 if goal is None else int(goal)
    par.flags = (FLAG_INIT_PER_REPLICA if per_replica_init else 0) | \
        (FLAG_ONE_SHOT if one_shot else 0)
    ip = None
    if init_perm is not None:
        ip = np.ascontiguousarray(init_perm, np.int32)
        par.init_perm = _p(ip, C.c_int32)
    else:
        par.init_perm = C.cast(None, _I32P)
    out = {
        "best_perm": np.zeros(n, np.int32),
        "final_perm": np.zeros((RT, n), np.int32),
        "final_E": np.zeros(RT, np.int64),
        "final_bestE": np.zeros(RT, np.int64),
        "final_pos": np.zeros(RT, np.int32),
        "final_accepted": np.zeros(RT, np.uint64),
        "best_perms": np.zeros((RT, n), np.int32),
    }
    megas = int(sweeps) // int(spe)
    traj_E = np.zeros((max(megas, 1), RT), np.int64) if traj else None
    res = _Result()
    res.best_perm = _p(out["best_perm"], C.c_int32)
    res.final_perm = _p(out["final_perm"], C.c_int32)
    res.final_E = _p(out["final_E"], C.c_int64)
    res.final_bestE = _p(out["final_bestE"], C.c_int64)
    res.final_pos = _p(out["final_pos"], C.c_int32)
    res.final_accepted = _p(out["final_accepted"], C.c_uint64)
    res.best_perms = _p(out["best_perms"], C.c_int32)
    if traj:
        res.traj_E = _p(traj_E, C.c_int64)
        res.traj_cap = megas
    else:
        res.traj_E = C.cast(None, _I64P)
        res.traj_cap = 0
    rc = lib().oracle_solve(C.byref(prob), C.byref(par), C.byref(res), int(mode), int(nthreads))
    if rc != 0:
        raise ValueError(f"oracle_solve failed with {rc}")
    out.update(best_cost=int(res.best_cost), attempts=int(res.attempts),
               accepted=int(res.accepted), exchanges=int(res.exchanges), kicks=int(res.kicks),
               megas_done=int(res.megas_done))
    if traj:
        out["traj_E"] = traj_E
    return out


def exchange_step(E, bestE, order, pos, tss, ladder_best, kick_after=100):
    """One exchange step (CPU_PIM.cpp:295-344) on copies; returns the new state."""
    E = np.ascontiguousarray(E, np.int64)
    bE = np.ascontiguousarray(bestE, np.int64)
    order = np.ascontiguousarray(order, np.int32).copy()
    pos = np.ascontiguousarray(pos, np.int32).copy()
    t = C.c_int32(int(tss))
    lb = C.c_int64(int(ladder_best))
    kicked = C.c_int32(0)
    nex = lib().oracle_exchange_step(E.size, _p(E, C.c_int64), _p(bE, C.c_int64),
                                     _p(order, C.c_int32), _p(pos, C.c_int32), C.byref(t),
                                     C.byref(lb), int(kick_after), C.byref(kicked))
    return {"order": order, "pos": pos, "tss": int(t.value), "ladder_best": int(lb.value),
            "exchanges": int(nex), "kicked": bool(kicked.value)}
