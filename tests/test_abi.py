"""The C-ABI library (libqap_pt.so): loads, exports every symbol include/qap_pt.h
declares, and validates arguments on the host before touching a device -- CPU only."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qap_pt.h")


def _declared():
    with open(HEADER) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(qap_(?:pt|srv)_\w+)\s*\(", txt)))


def test_header_declares_what_the_binding_lists(qpt):
    assert _declared() == sorted(qpt.EXPORTS)


def test_library_exports_every_declared_symbol(qpt):
    lib = qpt.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header(qpt):
    import ctypes as C
    # field order and widths of the public structs (x86-64 SysV layout)
    assert C.sizeof(qpt.QapProblem) == 24
    assert qpt.QapPtParams.init_perm.offset == C.sizeof(qpt.QapPtParams) - 8
    assert qpt.QapPtResult.seconds.offset == C.sizeof(qpt.QapPtResult) - 8


def _expect(qpt, code, inst):
    with pytest.raises(qpt.QapError) as ei:
        qpt.Pool([inst])
    assert ei.value.code == code, str(ei.value)


def test_host_validation_errors(qpt):
    F, D = qpt.instances.random_instance(1, 8)
    ok = dict(F=F, D=D, beta_ln=np.array([0.1, 0.2]), sweeps=4)
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(**dict(ok, beta_ln=np.array([0.2, 0.1]))))
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(**dict(ok, beta_ln=np.array([0.0, 0.1]))))
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(**dict(ok, sweeps=5, sweeps_per_exchange=2)))
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(**dict(ok, init_perm=np.zeros(8, np.int32))))
    # entries of 2^30: the deltas exceed even the int64 force matrices' bound
    big = np.full((8, 8), 1 << 30, dtype=np.int32)
    np.fill_diagonal(big, 0)
    _expect(qpt, qpt.QAP_PT_E_OVERFLOW, qpt.Instance(**dict(ok, F=big, D=big)))
    # entries of 2^20 pass the host checks (int64 force matrices); without a device the
    # pool then fails on the device, not on the bounds
    wide = np.full((8, 8), 1 << 20, dtype=np.int32)
    np.fill_diagonal(wide, 0)
    if qpt.device_count() < 1:
        _expect(qpt, qpt.QAP_PT_E_DEVICE, qpt.Instance(**dict(ok, F=wide, D=wide)))
    else:
        qpt.Pool([qpt.Instance(**dict(ok, F=wide, D=wide))]).close()
    Fz = np.zeros((QMAX := qpt.MAX_N + 1, QMAX), np.int32)
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(**dict(ok, F=Fz, D=Fz)))


def test_one_shot_limits_are_checked_on_the_host(qpt):
    """One-shot steps need n <= 256 and update vectors that split into two signed bytes
    (2 max|D|, 4 max|F| <= 32639); anything else is refused before any device work."""
    F, D = qpt.instances.random_instance(2, 257)
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(F=F, D=D, beta_ln=np.array([0.1]), sweeps=1,
                                                 one_shot=True))
    F, D = qpt.instances.random_instance(3, 8)
    F = F * 1000  # 4 max|F| = 36000
    _expect(qpt, qpt.QAP_PT_E_INVAL, qpt.Instance(F=F, D=D, beta_ln=np.array([0.1]), sweeps=1,
                                                 one_shot=True))
    import ctypes as C
    F, D = qpt.instances.random_instance(4, 8)
    inst = qpt.Instance(F=F, D=D, beta_ln=np.array([0.1]), sweeps=1)
    prob, par = inst.c_structs()
    par.flags = 8  # no such flag
    out = C.c_void_p()
    rc = qpt.load_library().qap_pt_pool_create(1, C.byref(prob), C.byref(par), C.byref(out))
    assert rc == qpt.QAP_PT_E_INVAL


def test_no_device_is_an_error_not_a_fallback(qpt):
    """Without a visible GPU a valid request fails with QAP_PT_E_DEVICE: there is no
    CPU path inside the product library."""
    if qpt.device_count() > 0:
        pytest.skip("a HIP device is visible")
    F, D = qpt.instances.random_instance(2, 8)
    _expect(qpt, qpt.QAP_PT_E_DEVICE, qpt.Instance(F=F, D=D, beta_ln=np.array([0.1]), sweeps=2))


def test_server_refuses_request_larger_than_budget(qpt):
    """Admission is gated on device memory: with an explicit budget a request whose
    footprint alone exceeds it is refused at submit (host side, no device needed)."""
    F, D = qpt.instances.random_instance(3, 64)
    inst = qpt.Instance(F=F, D=D, beta_ln=np.linspace(0.1, 1.0, 16), sweeps=4)
    with qpt.Server(device_budget=100_000) as srv:
        with pytest.raises(qpt.QapError) as ei:
            srv.submit(inst)
        assert ei.value.code == qpt.QAP_PT_E_NOMEM
        small = qpt.Instance(F=F[:8, :8].copy(), D=D[:8, :8].copy(), beta_ln=np.array([0.1]),
                             sweeps=2)
        t = srv.submit(small)
        r = srv.wait(t, timeout_s=30.0)
        if qpt.device_count() == 0:
            assert r is not None and r["status"] == qpt.QAP_PT_E_DEVICE
        else:
            assert r is not None and r["status"] == qpt.QAP_SRV_DONE


def test_server_shutdown_fails_queued_requests(qpt):
    F, D = qpt.instances.random_instance(4, 10)
    srv = qpt.Server(max_pools=1)
    for _ in range(3):
        srv.submit(qpt.Instance(F=F, D=D, beta_ln=np.array([0.1, 0.2]), sweeps=2))
    srv.close()  # must return: queued requests are failed, running pools finish


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures the one-GPU path only: --gpus above 1 exits with a message
    (status 2) before touching a device."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stderr
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
