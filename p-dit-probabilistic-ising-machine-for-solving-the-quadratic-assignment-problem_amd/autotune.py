"""Sweep-kernel shape autotuner: writes the tables libqap_pt.so reads at pool creation.

For each size bucket, flow class (dense: U[0,99]; sparse: density 0.1, U[1,99] -- the
batched generators) and force-matrix placement the size admits, every supported
shape (threads x candidates per thread) with n <= threads*cpl < 2n (at least 64x1) is
timed in one of two regimes:
  --regime full     (default) a full-chip load: one instance with enough ladders of 32
                    replicas to give `--replicas` (4096) workgroups, HIP-event time of the
                    sweep launches -> tune/shapes_gfx950.txt;
  --regime latency  one replica per CU at most: `--replicas` (64) workgroups, wall time of
                    the launches such a pool runs (persistent launches for the in-LDS
                    placements, graph replays otherwise) -> tune/shapes_latency_gfx950.txt.
One exchange per sweep.  The fastest shape per (bucket, flow, placement) becomes a row
    n_max  dense|sparse  lds32|lds16|hbm|lazy  nt  cpl  rate
with every measured rate in a comment line above it.  Each placement is forced with
QAP_PT_IMODE; up to n=256 every placement is timed on the same binary-flow data (which
the uint16 placement admits), so the rates compare placements and the library picks the
fastest placement that fits an instance.

    python <package>/autotune.py [--regime full|latency] [--sizes 32,64,...]
                                 [--replicas N] [--out PATH]
"""
from __future__ import annotations

import argparse
import importlib.util
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = {"full": os.path.join(HERE, "tune", "shapes_gfx950.txt"),
               "latency": os.path.join(HERE, "tune", "shapes_latency_gfx950.txt")}
SHAPES = [(64, 1), (64, 2), (64, 4), (128, 1), (128, 2), (256, 1), (256, 2), (256, 4)]
LAZY_SHAPES = [(512, 1), (512, 2)]  # eight-wave workgroups exist for the lazy kernel only
PLACEMENTS = {0: "lds32", 1: "lds16", 2: "hbm", 3: "lazy"}
DEFAULT_SIZES = [32, 48, 64, 96, 128, 160, 192, 256, 384, 512, 768, 1024]


def _load_pkg():
    if "qap_pt_amd" in sys.modules:
        return sys.modules["qap_pt_amd"]
    spec = importlib.util.spec_from_file_location(
        "qap_pt_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["qap_pt_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def candidate_shapes(n: int, placement: int = 0):
    shapes = SHAPES + (LAZY_SHAPES if placement == 3 else [])
    c = [(nt, cpl) for nt, cpl in shapes if n <= nt * cpl < 2 * n]
    return c or [(64, 1)]


def _instance(pkg, n: int, dense: bool, ladders: int):
    F, D = pkg.instances.mixed_instance(0x7A11 + n, n, dense)
    if n <= 256:
        F = (F != 0).astype(F.dtype)  # binary flow keeps the force bound in uint16
    betas = pkg.instances.geometric_ladder(32, 1.0, 100.0) / max(1.0, n / 16.0)
    return pkg.Instance(F=F, D=D, beta_ln=betas, sweeps=0, T=ladders, seed=11)


def measure(pkg, n, dense, placement, shape, replicas, min_attempts, latency=False):
    """Swap attempts per second for one forced shape and placement -- of kernel time at
    full load, of wall time through the production launches when latency-bound -- or None
    when the library does not place this instance as asked."""
    os.environ["QAP_PT_IMODE"] = str(placement)
    os.environ["QAP_PT_SHAPE"] = f"{shape[0]},{shape[1]}"
    try:
        inst = _instance(pkg, n, dense, max(1, replicas // 32))
        with pkg.Pool([inst]) as pool:
            got = pool.shape(0)
            want = f"{shape[0]}x{shape[1]},{PLACEMENTS[placement]},"
            if not got.startswith(want):
                return None, got
            sweeps = max(2, math.ceil(min_attempts / (inst.R * inst.T * n)))
            pool.run(1)
            before = pool.result(0)["attempts"]
            if latency:
                t0 = time.perf_counter()
                pool.run(sweeps)
                secs = time.perf_counter() - t0
            else:
                ms, _ = pool.run(sweeps, timed=True)
                secs = ms / 1e3
            att = pool.result(0)["attempts"] - before
        return att / secs, got
    finally:
        os.environ.pop("QAP_PT_IMODE", None)
        os.environ.pop("QAP_PT_SHAPE", None)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--regime", default="full", choices=["full", "latency"])
    ap.add_argument("--sizes", default=",".join(map(str, DEFAULT_SIZES)))
    ap.add_argument("--replicas", type=int, default=0,
                    help="workgroups per run (default 4096 full, 64 latency)")
    ap.add_argument("--min-attempts", type=float, default=0,
                    help="attempts timed per measurement (default 2e7 full, 4e6 latency)")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    latency = args.regime == "latency"
    args.replicas = args.replicas or (64 if latency else 4096)
    args.min_attempts = args.min_attempts or (4e6 if latency else 2e7)
    args.out = args.out or DEFAULT_OUT[args.regime]
    # time the kernels, not an earlier table
    os.environ["QAP_PT_TUNE_TABLE"] = ""
    os.environ["QAP_PT_TUNE_LATENCY_TABLE"] = ""
    pkg = _load_pkg()
    if pkg.device_count() < 1:
        raise SystemExit("autotune: no HIP device")
    sizes = sorted(int(x) for x in args.sizes.split(","))
    what = ("wall time of the production launches (latency-bound: at most one replica per CU)"
            if latency else "kernel time")
    lines = [f"# sweep-kernel shapes for gfx950 ({args.regime} regime), written by autotune.py",
             f"# load: {args.replicas} replicas (ladders of 32) per run, >= "
             f"{args.min_attempts:.0e} attempts timed; rates are swap-attempts/s of {what}",
             "# n_max flow placement nt cpl rate"]
    t0 = time.time()
    for n in sizes:
        for dense in (True, False):
            flow = "dense" if dense else "sparse"
            for placement in (0, 1, 2, 3):
                rates = []
                for shape in candidate_shapes(n, placement):
                    rate, got = measure(pkg, n, dense, placement, shape, args.replicas,
                                        args.min_attempts, latency)
                    print(f"[{time.time() - t0:7.1f}s] n={n} {flow} {PLACEMENTS[placement]} "
                          f"{shape[0]}x{shape[1]}: "
                          + (f"{rate:.4g}/s" if rate else f"not placed ({got})"), flush=True)
                    if rate:
                        rates.append((rate, shape))
                if not rates:
                    continue
                rates.sort(reverse=True)
                rate, best = rates[0]
                lines.append("#   " + "  ".join(f"{s[0]}x{s[1]} {r:.4g}" for r, s in rates))
                lines.append(f"{n} {flow} {PLACEMENTS[placement]} {best[0]} {best[1]} {rate:.4g}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
