"""Automatic ladder rule and the command-line front end, host side (no device).

The rule (CPU_PIM.cpp:431-473) is checked on average-energy curves whose expected
ladder is worked out by hand in each test's comment; the command line is checked for
its argument handling and for failing loudly without a device.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR

CLI = os.path.join(PKG_DIR, "lib", "qap_pt_cli")


def _curve(fn):
    return np.array([fn(j) for j in range(32)], dtype=np.float64)


def test_auto_select_linear_descent(qpt):
    # avg = 1000 up to exponent index 10, then falls by 90 per step to 100 at index 20 and
    # stays.  maxE = avg[0] = 1000, minE = 100 first at index 20.  The minimum beta is the
    # first index below 0.9*900+100 = 910: index 12 (avg 820) -> 2^(-30+12).  The maximum
    # is two past the first later index below 0.1*900+100 = 190 or at the minimum: index
    # 20 -> 2^(-30+22).  ratio = (2^-8 / 2^-18)^(1/R).
    avg = _curve(lambda j: 1000.0 if j <= 10 else max(100.0, 1000.0 - 90.0 * (j - 10)))
    R = 8
    b0, m = qpt.auto_select(avg, R)
    assert b0 == pytest.approx(2.0 ** -8 / 1.44269504089, rel=1e-15)
    assert m == pytest.approx(2.0 ** (10.0 / R), rel=1e-14)


def test_auto_select_isolated_dip_uses_index_before_minimum(qpt):
    # Flat at 1000 except 0 at index 20: no index but 20 is below 910, so the minimum beta
    # comes from the "one before the minimum" clause (index 19), and the maximum from the
    # minimum itself (index 20) -> 2^(-30+22); ratio = (2^-8 / 2^-11)^(1/R).
    avg = _curve(lambda j: 0.0 if j == 20 else 1000.0)
    b0, m = qpt.auto_select(avg, 10)
    assert b0 == pytest.approx(2.0 ** -8 / 1.44269504089, rel=1e-15)
    assert m == pytest.approx(2.0 ** (3.0 / 10), rel=1e-14)


def test_auto_select_flat_curve_spans_whole_scan(qpt):
    # A flat curve never crosses either threshold and its minimum is index 0: the minimum
    # beta falls back to 2^-30 and the maximum is two past the last index considered
    # (30) -> 2^2.  ratio = 2^(32/R).
    avg = _curve(lambda j: 500.0)
    b0, m = qpt.auto_select(avg, 4)
    assert b0 == pytest.approx(4.0 / 1.44269504089, rel=1e-15)
    assert m == pytest.approx(2.0 ** 8, rel=1e-13)


def test_auto_select_ignores_last_exponent(qpt):
    # The last exponent's average never enters the rule (the reference evaluates it with
    # j < i at i = 31): a huge drop there changes nothing.
    base = _curve(lambda j: 1000.0 if j <= 10 else max(100.0, 1000.0 - 90.0 * (j - 10)))
    moved = base.copy()
    moved[31] = -1e9
    assert qpt.auto_select(base, 8) == qpt.auto_select(moved, 8)


def test_auto_select_rejects_bad_arguments(qpt):
    with pytest.raises(qpt.QapError) as e:
        qpt.auto_select(np.zeros(1), 4)
    assert e.value.code == qpt.QAP_PT_E_INVAL
    with pytest.raises(qpt.QapError):
        qpt.auto_select(np.zeros(33), 4)
    with pytest.raises(qpt.QapError):
        qpt.auto_select(np.zeros(32), 0)


def test_ladder_from_matches_reference_order(qpt):
    # CPU_PIM.cpp:113-118: the last replica holds beta0 and each hotter one is the next
    # colder divided by the ratio.
    lad = qpt.ladder_from(0.5, 2.0, 4)
    np.testing.assert_array_equal(lad, [0.0625, 0.125, 0.25, 0.5])
    assert math.isclose(qpt.ladder_from(1.0, 1.0, 3)[0], 1.0)


def _write_dat(path, F, D):
    n = F.shape[0]
    with open(path, "w") as fh:
        fh.write(f"{n}\n\n")
        for M in (F, D):
            for row in M:
                fh.write(" ".join(str(int(x)) for x in row) + "\n")
            fh.write("\n")


def test_cli_argument_errors(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "No input file" in r.stderr
    r = subprocess.run([CLI, "-r"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "requires a value" in r.stderr
    r = subprocess.run([CLI, "-r", "4", str(tmp_path / "missing.dat")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "Could not open" in r.stderr
    bad = tmp_path / "short.dat"
    bad.write_text("3\n1 2 3\n")
    r = subprocess.run([CLI, "-b", "1", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ends before" in r.stderr


def test_cli_without_device_fails_loudly(tmp_path, qpt):
    if qpt.device_count() > 0:
        pytest.skip("a HIP device is visible")
    F = np.arange(9, dtype=np.int32).reshape(3, 3)
    path = tmp_path / "t3.dat"
    _write_dat(path, F, F.T.copy())
    r = subprocess.run([CLI, "-b", "1", "-r", "2", "-w", "2", str(path)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr
