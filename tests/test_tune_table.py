"""The tuned shape tables: parsing and precedence (host only, no device).

The library reads the tables once per process, so each case runs in a fresh Python
process with QAP_PT_TUNE_TABLE (full-chip) and QAP_PT_TUNE_LATENCY_TABLE (latency-bound
pools) pointing at files written here, or empty.
"""
import os
import subprocess
import sys
import textwrap

from conftest import ROOT

PROBE = textwrap.dedent("""
    import sys
    sys.path.insert(0, {tests!r})
    from conftest import load_qpt
    q = load_qpt()
    for n in (16, 40, 64, 65, 100, 129, 200, 300, 700):
        print(n, "lat", q.kernel_shape(n))
        print(n, "thr", q.kernel_shape(n, 4096))
""")


def _shapes(tmp_path, table_text, extra_env=None, latency_text=None):
    env = dict(os.environ)
    env.pop("QAP_PT_SHAPE", None)
    env.pop("QAP_PT_IMODE", None)
    for var, text, name in (("QAP_PT_TUNE_TABLE", table_text, "shapes.txt"),
                            ("QAP_PT_TUNE_LATENCY_TABLE", latency_text, "latency.txt")):
        if text is None:
            env[var] = ""
        else:
            path = tmp_path / name
            path.write_text(text)
            env[var] = str(path)
    env.update(extra_env or {})
    out = subprocess.run([sys.executable, "-c", PROBE.format(tests=os.path.join(ROOT, "tests"))],
                         capture_output=True, text=True, env=env, timeout=120, check=True).stdout
    shapes = {}
    for line in out.strip().splitlines():
        n, regime, shape = line.split(" ", 2)
        shapes[n if regime == "lat" else n + "@full"] = shape
    return shapes


def test_size_rule_without_table(tmp_path):
    s = _shapes(tmp_path, None)
    assert s["16"].startswith("64x1,lds32")
    assert s["65"].startswith("128x1,lds32")
    assert s["200"].startswith("256x1,")
    assert s["300"].startswith("512x1,lazy")   # symmetric, too big for LDS: eight waves
    assert s["700"].startswith("512x2,lazy")


def test_table_rows_apply_by_size_flow_and_placement(tmp_path):
    table = textwrap.dedent("""\
        # comment line
        64 dense lds32 64 1   # trailing comment
        128 dense lds32 64 2
        256 dense hbm 128 2
        1024 dense lazy 256 4
        128 sparse lds32 128 1
        """)
    s = _shapes(tmp_path, table)
    assert s["40"].startswith("64x1,lds32")
    assert s["100"].startswith("64x2,lds32")      # second dense lds32 row
    assert s["129"].startswith("256x1,")          # lds32 row ends at 128; size rule
    assert s["300"].startswith("256x4,lazy")      # 1024-row of the lazy placement
    assert s["700"].startswith("256x4,lazy")


def test_bad_rows_are_ignored_and_shapes_must_cover_n(tmp_path):
    table = textwrap.dedent("""\
        64 dense lds32 96 1
        64 dense lds32
        garbage
        128 dense lds32 64 1
        200 dense lds64 64 4
        """)
    s = _shapes(tmp_path, table)
    assert s["40"].startswith("64x1,lds32")      # 96x1 is not a kernel shape
    assert s["100"].startswith("128x1,lds32")    # 64x1 cannot hold 100 candidates
    assert s["200"].startswith("256x1,")         # unknown placement name


def test_environment_shape_overrides_table(tmp_path):
    table = "128 dense lds32 64 2\n"
    s = _shapes(tmp_path, table, {"QAP_PT_SHAPE": "128,1"})
    assert s["100"].startswith("128x1,lds32")


def test_rates_choose_among_placements_that_fit(tmp_path):
    # n=64 fits the int32 LDS placement (and the lazy one: the probe is symmetric).  In a
    # pool with more replicas than CUs, a table that rates the in-HBM placement highest
    # for n <= 64 moves it there, while n=100 (no hbm or lazy row rated) keeps the default
    # order; a small (latency-bound) pool keeps the LDS placement either way.
    table = textwrap.dedent("""\
        64 dense lds32 64 1 1.0e8
        64 dense lazy 64 1 0.5e8
        64 dense hbm 64 1 2.0e8
        128 dense lds32 128 1 1.0e8
        """)
    s = _shapes(tmp_path, table)
    assert s["40@full"].startswith("64x1,hbm")
    assert s["100@full"].startswith("128x1,lds32")
    assert s["40"].startswith("64x1,lds32")
    # an explicit placement request still wins
    s = _shapes(tmp_path, table, {"QAP_PT_IMODE": "0"})
    assert s["40@full"].startswith("64x1,lds32")


def test_latency_table_sets_shapes_of_small_pools_only(tmp_path):
    # a latency-bound pool (at most one replica per CU) takes its shape from the latency
    # rows, falling back to the full-chip rows; its placement keeps the fixed order (the
    # latency rates do not move it); a full-chip pool ignores the latency rows
    latency = textwrap.dedent("""\
        64 dense lds32 64 1 2.0e7
        64 dense lazy 64 1 3.0e7
        128 dense lds32 64 2 1.5e7
        """)
    full = "256 dense lds32 128 2\n"
    s = _shapes(tmp_path, full, latency_text=latency)
    assert s["40"].startswith("64x1,lds32")       # placement order, not the lazy rate
    assert s["100"].startswith("64x2,lds32")      # the latency row's shape
    assert s["129"].startswith("128x2,lds32")     # no latency row: the full-chip row
    assert s["100@full"].startswith("128x2,lds32")  # full-chip pools: full-chip rows only
