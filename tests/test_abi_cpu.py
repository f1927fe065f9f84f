"""The C ABI without a GPU: the library loads, exports every symbol include/qsim.h
declares, and rejects bad arguments with a status and a message (no device call)."""

import ctypes
import os
import re

import pytest

from qsim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, 'include', 'qsim.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(qs_[a-z0-9_]+)\s*\(', src)))


def test_library_exports_every_declared_function():
    lib = _lib.load_library()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python prototypes cover the header
    assert set(names) <= set(_lib._PROTOS), set(names) - set(_lib._PROTOS)
    assert _lib.missing_symbols() == []


def test_abi_version_and_limits():
    lib = _lib.load_library()
    assert lib.qs_abi_version() == 2
    assert lib.qs_sv_max_qubits() == 34
    assert lib.qs_mps_max_chi() == 128
    ntasks, chi = 5, 64
    ws = lib.qs_mps_workspace_bytes(ntasks, chi)
    theta = ntasks * 4 * chi * chi * 16          # one (2chi x 2chi) complex128 slot per task
    assert theta < ws < theta + ntasks * 1024 + 256
    # chi_max above 64: theta plus the SVD's working copy of its columns
    ws128 = lib.qs_mps_workspace_bytes(ntasks, 128)
    assert ntasks * 8 * 128 * 128 * 16 < ws128 < ntasks * 8 * 128 * 128 * 16 + ntasks * 1024 + 256
    assert lib.qs_mps_workspace_bytes(0, chi) == 0
    assert ctypes.sizeof(_lib.QsGate) == 16 + 8 + 256
    assert ctypes.sizeof(_lib.QsMpsTask) == 8 * 3 + 8 + 4 + 4 + 8 + 4 + 4 + 256


def test_svd_team_width_table():
    """The SVD tile-shape table the autotuner drives: choices, set/get, rejection."""
    from qsim_amd import mps_tune
    lib = _lib.load_library()
    expected = {8: [4, 8, 16], 16: [4, 8, 16, 32], 32: [4, 8, 16, 32], 64: [8, 16]}
    saved = {c: mps_tune.current_lanes(c) for c in expected}
    try:
        for c, choices in expected.items():
            assert mps_tune.lane_choices(c) == choices
            for w in choices:                       # threads per task = w * c <= 1024
                assert w * c <= 1024 and 2 * c // w <= 16
                mps_tune.set_lanes(c, w)
                assert mps_tune.current_lanes(c) == w
        assert lib.qs_mps_set_svd_lanes(64, 4) == _lib.QS_E_ARG      # not built
        assert lib.qs_mps_set_svd_lanes(128, 16) == _lib.QS_E_ARG    # workspace class
        assert lib.qs_mps_svd_lanes(128) == _lib.QS_E_ARG
        cnt = ctypes.c_int(-1)
        assert lib.qs_mps_svd_lane_choices(12, None, 0, ctypes.byref(cnt)) == _lib.QS_E_ARG
    finally:
        for c, w in saved.items():
            mps_tune.set_lanes(c, w)


def test_batch_plan_groups_by_svd_class_without_device():
    """qs_mps_batch_plan runs on the host: class grouping, original indices, errors."""
    import numpy as np
    lib = _lib.load_library()
    chis = [8, 64, 16, 128, 8, 32, 100]
    tasks = (_lib.QsMpsTask * len(chis))()
    for k, c in enumerate(chis):
        t = tasks[k]
        t.sites, t.bond, t.dlog = 0x10000, 0x20000, 0x30000 + 8 * k
        t.site_stride = 2 * 128 * 128
        t.site = 3 * k
        t.n_sites = 3 * len(chis)
        t.chi_max = c
        t.cutoff = 1e-10
    nb = lib.qs_mps_batch_table_bytes(len(chis))
    rec = nb // len(chis)
    assert rec >= ctypes.sizeof(_lib.QsMpsTask) + 4
    buf = np.zeros(nb, dtype=np.uint8)
    counts = (ctypes.c_int32 * 5)()
    chi = ctypes.c_int(0)
    st = lib.qs_mps_batch_plan(tasks, len(chis), ctypes.c_void_p(buf.ctypes.data), nb, counts,
                               ctypes.byref(chi))
    assert st == 0, lib.qs_last_error()
    assert list(counts) == [2, 1, 1, 1, 2] and chi.value == 128
    size = ctypes.sizeof(_lib.QsMpsTask)
    orig = [int(np.frombuffer(buf[k * rec + size:k * rec + size + 4].tobytes(), np.int32)[0])
            for k in range(len(chis))]
    assert orig == [0, 4, 2, 5, 1, 3, 6]
    # the table holds each task verbatim
    for k, o in enumerate(orig):
        assert buf[k * rec:k * rec + size].tobytes() == bytes(tasks[o])
    assert lib.qs_mps_batch_plan(tasks, len(chis), ctypes.c_void_p(buf.ctypes.data), nb - 1,
                                 counts, ctypes.byref(chi)) == _lib.QS_E_RANGE
    tasks[6].site = tasks[6].n_sites - 1                    # bond (n-1, n) past the chain
    assert lib.qs_mps_batch_plan(tasks, len(chis), ctypes.c_void_p(buf.ctypes.data), nb, counts,
                                 ctypes.byref(chi)) == _lib.QS_E_ARG
    assert b'outside the chain' in lib.qs_last_error()
    tasks[6].site = 18
    tasks[6].n_sites = 0                                    # chain length missing
    assert lib.qs_mps_batch_plan(tasks, len(chis), ctypes.c_void_p(buf.ctypes.data), nb, counts,
                                 ctypes.byref(chi)) == _lib.QS_E_ARG
    tasks[6].n_sites = 3 * len(chis)
    tasks[1].site = tasks[0].site + 1                       # overlapping sites
    assert lib.qs_mps_batch_plan(tasks, len(chis), ctypes.c_void_p(buf.ctypes.data), nb, counts,
                                 ctypes.byref(chi)) == _lib.QS_E_ARG
    # launch-side argument checks come before any device work
    bad = (ctypes.c_int32 * 5)(1, 0, 0, 0, 0)
    ws = ctypes.c_void_p(0x40000)
    assert lib.qs_mps_batch_launch(ctypes.c_void_p(0x50000), 2, bad, 8, 0, ws, 1 << 30, None,
                                   None) == _lib.QS_E_ARG        # counts sum to 1, not 2
    hi = (ctypes.c_int32 * 5)(0, 0, 0, 0, 2)
    assert lib.qs_mps_batch_launch(ctypes.c_void_p(0x50000), 2, hi, 64, 0, ws, 1 << 30, None,
                                   None) == _lib.QS_E_ARG        # class 128 above chi_max 64
    ok = (ctypes.c_int32 * 5)(2, 0, 0, 0, 0)
    assert lib.qs_mps_batch_launch(ctypes.c_void_p(0x50000), 2, ok, 8, 0, ws, 16, None,
                                   None) == _lib.QS_E_RANGE      # workspace too small


def test_argument_errors_are_reported_without_device():
    lib = _lib.load_library()
    g = (_lib.QsGate * 1)()
    g[0].kind = _lib.QS_GATE_U1
    g[0].q[0] = 5
    n = ctypes.c_int(0)
    st = lib.qs_sv_plan_passes(3, g, 1, ctypes.byref(n))
    assert st == _lib.QS_E_ARG
    assert b'out of range' in lib.qs_last_error()
    st = lib.qs_sv_plan_passes(40, g, 1, ctypes.byref(n))
    assert st == _lib.QS_E_RANGE
    g[0].q[0] = 0
    g[0].ctrl_mask = 1
    assert lib.qs_sv_plan_passes(3, g, 1, ctypes.byref(n)) == _lib.QS_E_ARG
    assert b'overlaps' in lib.qs_last_error()
    g[0].ctrl_mask = 0
    assert lib.qs_sv_plan_passes(3, g, 1, ctypes.byref(n)) == _lib.QS_OK and n.value == 1
    assert lib.qs_last_error() == b''
    # null state pointer
    assert lib.qs_sv_set_zero(None, 3, None) == _lib.QS_E_ARG
    # plan objects: argument errors before any device work
    h = ctypes.c_void_p()
    assert lib.qs_sv_plan_create(40, g, 1, ctypes.byref(h)) == _lib.QS_E_RANGE and not h.value
    assert lib.qs_sv_plan_create(3, None, 1, ctypes.byref(h)) == _lib.QS_E_ARG
    assert lib.qs_sv_plan_create(3, g, 1, None) == _lib.QS_E_ARG
    assert lib.qs_sv_plan_launch(None, ctypes.c_void_p(16), None) == _lib.QS_E_ARG
    assert lib.qs_sv_plan_info(None, None, None, None) == _lib.QS_E_ARG
    assert lib.qs_sv_plan_destroy(None) == _lib.QS_OK
    # measurement shots: workspace size, null state, negative shot count
    assert lib.qs_sv_sample_workspace_bytes(12) == 2 * 8
    assert lib.qs_sv_sample_workspace_bytes(28) == ((1 << 16) + 1) * 8
    assert lib.qs_sv_sample_workspace_bytes(0) == 0
    assert lib.qs_sv_sample(None, 3, None, 4, None, None, 0, None) == _lib.QS_E_ARG
    assert lib.qs_sv_sample(ctypes.c_void_p(16), 3, None, -1, None, None, 0, None) == _lib.QS_E_ARG
    assert b'negative shot count' in lib.qs_last_error()
    # MPS: chi beyond the limit is an error, never a clamp
    t = (_lib.QsMpsTask * 1)()
    t[0].sites = 1
    t[0].bond = 1
    t[0].dlog = 1
    t[0].n_sites = 4
    t[0].chi_max = 129
    t[0].site_stride = 2 * 129 * 129
    assert lib.qs_mps_update_batch(t, 1, 0, ctypes.c_void_p(1), 1 << 30, None, None) == _lib.QS_E_CHI
    t[0].chi_max = 8
    assert lib.qs_mps_update_batch(t, 1, 7, ctypes.c_void_p(1), 1 << 30, None, None) == _lib.QS_E_ARG
    assert lib.qs_mps_update_batch(t, 1, 0, ctypes.c_void_p(1), 16, None, None) == _lib.QS_E_RANGE
    t2 = (_lib.QsMpsTask * 2)()
    for k in range(2):
        t2[k].sites, t2[k].bond, t2[k].dlog = 1, 1, 1
        t2[k].chi_max, t2[k].site_stride, t2[k].site = 8, 128, 3 + k
        t2[k].n_sites = 8
    assert lib.qs_mps_update_batch(t2, 2, 0, ctypes.c_void_p(1), 1 << 30, None, None) == _lib.QS_E_ARG
    assert b'share a site' in lib.qs_last_error()
    # entanglement entropy: the bond list is checked on the host
    bl = (ctypes.c_int32 * 2)(0, 7)
    one = ctypes.c_void_p(16)
    assert lib.qs_mps_entanglement_entropy(one, 8, 16, one, one, bl, 2, one, one, 1 << 20,
                                           None) == _lib.QS_E_ARG
    assert b'out of range' in lib.qs_last_error()
    # batched one-site gates and site resets: argument checks before any launch
    sl = (ctypes.c_int32 * 3)(1, 4, 1)
    mats = (ctypes.c_double * 24)()
    one = ctypes.c_void_p(16)
    assert lib.qs_mps_apply_1q_batch(one, one, 128, 8, sl, mats, 3, None) == _lib.QS_E_ARG
    assert b'appears twice' in lib.qs_last_error()
    sl[2] = 8
    assert lib.qs_mps_apply_1q_batch(one, one, 128, 8, sl, mats, 3, None) == _lib.QS_E_ARG
    assert b'outside the chain' in lib.qs_last_error()
    assert lib.qs_mps_apply_1q_batch(one, one, 128, 8, None, None, 0, None) == _lib.QS_OK
    assert lib.qs_mps_reset_site(one, one, 128, 8, 8, None) == _lib.QS_E_ARG
    assert lib.qs_mps_reset_site(None, one, 128, 8, 0, None) == _lib.QS_E_ARG
    # the allocation-free batch: argument checks and the scratch limit come first
    assert lib.qs_mps_update_batch_static(t, 1, 7, None, None) == _lib.QS_E_ARG
    assert lib.qs_mps_update_batch_static(t2, 2, 0, None, None) == _lib.QS_E_ARG
    n = 130
    big = (_lib.QsMpsTask * n)()
    for k in range(n):
        big[k].sites, big[k].bond, big[k].dlog = 1, 1, 1
        big[k].chi_max, big[k].site_stride, big[k].site = 64, 2 * 64 * 64, 2 * k
        big[k].n_sites = 2 * n + 1
    assert lib.qs_mps_update_batch_static(big, n, 1, None, None) == _lib.QS_E_RANGE
    assert b'static scratch' in lib.qs_last_error()
    need = lib.qs_mps_workspace_bytes(128, 64)
    assert need > _lib.QS_MPS_STATIC_SCRATCH_BYTES >= lib.qs_mps_workspace_bytes(127, 64)


def test_engines_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip('a GPU is visible')
    from qsim_amd.sv import StateVector
    with pytest.raises(RuntimeError):
        StateVector(3)
    from fpga_mps_simulator import FPGAMPSSimulator
    with pytest.raises(RuntimeError):
        FPGAMPSSimulator(chi_max=8)
    with pytest.raises(ValueError):
        FPGAMPSSimulator(num_cards=2)
