"""Synthetic QAP instances and temperature ladders for the benchmark workloads.

Every generator is deterministic in its seed (numpy's Philox bit generator) so the
same instance can be rebuilt on any machine.  Matrices are int32, row-major, flow
first, as the reference reads them (CPU-PIM/CPU_PIM.cpp:79-107).
"""
from __future__ import annotations

import math

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(int(seed)))


def geometric_ladder(R: int, t_min: float, t_max: float) -> np.ndarray:
    """beta (natural-log units) of R rungs, temperatures geometric from t_max (rung 0,
    hottest) down to t_min (rung R-1); ascending as the reference orders them
    (CPU_PIM.cpp:113-118: beta_l = beta0 / m^(R-1-l))."""
    if R == 1:
        return np.array([1.0 / t_min])
    k = np.arange(R, dtype=np.float64)
    temps = t_max * (t_min / t_max) ** (k / (R - 1))
    return 1.0 / temps


def beta_ladder_from_flags(R: int, beta0_ln: float, ratio: float) -> np.ndarray:
    """The reference's -b/-m ladder: largest beta beta0 (ln units), each rung below
    divided by `ratio` in turn (CPU_PIM.cpp:111-119)."""
    out = np.zeros(R, dtype=np.float64)
    b = float(beta0_ln)
    for l in range(R - 1, -1, -1):
        out[l] = b
        b /= ratio
    return out


def _sym_uniform(g: np.random.Generator, n: int, lo: int, hi: int) -> np.ndarray:
    up = g.integers(lo, hi + 1, size=(n, n), dtype=np.int64)
    F = np.triu(up, 1)
    F = F + F.T
    return F.astype(np.int32)


def _sym_sparse(g: np.random.Generator, n: int, density: float, lo: int, hi: int) -> np.ndarray:
    vals = g.integers(lo, hi + 1, size=(n, n), dtype=np.int64)
    mask = g.random((n, n)) < density
    F = np.triu(np.where(mask, vals, 0), 1)
    F = F + F.T
    return F.astype(np.int32)


def euclid_dist(g: np.random.Generator, n: int, side: float = 100.0) -> np.ndarray:
    """Rounded (half away from zero) Euclidean distance of n points uniform in a square."""
    pts = g.random((n, 2)) * side
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    return np.floor(d + 0.5).astype(np.int32)


def flat_c1(seed: int = 0xC0FFEE + 1, n: int = 64):
    """n=64: symmetric zero-diagonal flow U[0,99]; rounded Euclidean distance of points
    in a 100x100 square."""
    g = rng(seed)
    return _sym_uniform(g, n, 0, 99), euclid_dist(g, n)


def grey_c2(n: int = 256, block: int = 92, side: int = 16):
    """Grey-pattern-shaped instance: F = ones on a block x block block (zero diagonal),
    D = squared wrapped Manhattan distance on a side x side torus."""
    assert side * side == n
    F = np.zeros((n, n), dtype=np.int32)
    F[:block, :block] = 1
    np.fill_diagonal(F, 0)
    xy = np.array([(k // side, k % side) for k in range(n)])
    dx = np.abs(xy[:, None, 0] - xy[None, :, 0])
    dy = np.abs(xy[:, None, 1] - xy[None, :, 1])
    dx = np.minimum(dx, side - dx)
    dy = np.minimum(dy, side - dy)
    D = ((dx + dy) ** 2).astype(np.int32)
    return F, D


def mixed_instance(seed: int, n: int, dense: bool):
    """Batched/served instance: dense flow U[0,99] or sparse flow (density 0.1,
    nonzeros U[1,99]), symmetric with zero diagonal; Euclidean distance."""
    g = rng(seed)
    F = _sym_uniform(g, n, 0, 99) if dense else _sym_sparse(g, n, 0.1, 1, 99)
    return F, euclid_dist(g, n)


def batch_c3(count: int = 512, seed: int = 0xC0FFEE + 3, n_lo: int = 30, n_hi: int = 150):
    """count instances, n uniform in [n_lo, n_hi], alternating dense / sparse flow."""
    g = rng(seed)
    ns = g.integers(n_lo, n_hi + 1, size=count)
    return [mixed_instance(seed * 1000 + k, int(ns[k]), dense=(k % 2 == 0))
            for k in range(count)]


def served_sizes(count: int, seed: int = 0xC0FFEE + 4, n_lo: int = 20, n_hi: int = 256):
    """n log-uniform in [n_lo, n_hi]."""
    g = rng(seed)
    u = g.random(count)
    return np.floor(np.exp(math.log(n_lo) + u * (math.log(n_hi + 1) - math.log(n_lo)))).astype(int)


def grid_manhattan(nx: int, ny: int) -> np.ndarray:
    xy = np.array([(k // ny, k % ny) for k in range(nx * ny)])
    d = np.abs(xy[:, None, :] - xy[None, :, :]).sum(-1)
    return d.astype(np.int32)


def large_c5(n: int = 512, seed: int = 0xC0FFEE + 5):
    """n=512 (32x16 grid) or n=1024 (32x32 grid): flow U[0,255] symmetric with zero
    diagonal; Manhattan distance on the grid."""
    g = rng(seed)
    ny = n // 32
    return _sym_uniform(g, n, 0, 255), grid_manhattan(32, ny)


def random_instance(seed: int, n: int, lo: int = 0, hi: int = 9, symmetric: bool = True,
                    zero_diag: bool = True):
    """Small random instance for parity tests (asymmetric / nonzero diagonal allowed)."""
    g = rng(seed)
    if symmetric:
        F = _sym_uniform(g, n, lo, hi)
        D = _sym_uniform(g, n, lo, hi)
    else:
        F = g.integers(lo, hi + 1, size=(n, n)).astype(np.int32)
        D = g.integers(lo, hi + 1, size=(n, n)).astype(np.int32)
    if zero_diag:
        np.fill_diagonal(F, 0)
        np.fill_diagonal(D, 0)
    elif symmetric:
        F[np.diag_indices(n)] = g.integers(lo, hi + 1, size=n)
        D[np.diag_indices(n)] = g.integers(lo, hi + 1, size=n)
    return F, D


def qap_cost(F: np.ndarray, D: np.ndarray, perm: np.ndarray) -> int:
    """E = sum_ij D[i][j] F[p(i)][p(j)] (CPU_PIM.cpp:184-192), exact in int64."""
    p = np.asarray(perm)
    return int((D.astype(np.int64) * F.astype(np.int64)[np.ix_(p, p)]).sum())


def read_qaplib(path: str):
    """readProblem (CPU_PIM.cpp:79-107): whitespace-separated n, then F (n*n), then D."""
    with open(path) as fh:
        toks = fh.read().split()
    n = int(toks[0])
    vals = np.array([int(t) for t in toks[1:1 + 2 * n * n]], dtype=np.int64)
    if vals.size != 2 * n * n:
        raise ValueError(f"{path}: expected {2 * n * n} matrix entries, found {vals.size}")
    F = vals[:n * n].reshape(n, n).astype(np.int32)
    D = vals[n * n:].reshape(n, n).astype(np.int32)
    return F, D


def write_qaplib(path: str, F: np.ndarray, D: np.ndarray):
    n = F.shape[0]
    with open(path, "w") as fh:
        fh.write(f"{n}\n\n")
        for M in (F, D):
            for row in M:
                fh.write(" ".join(str(int(v)) for v in row) + "\n")
            fh.write("\n")
