"""ctypes binding of libqsim_hip.so (declarations in include/qsim.h).

The library is the only compute path: when it is missing, or no gfx950 device
is visible, every engine constructor raises RuntimeError.  There is no CPU
fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, 'libqsim_hip.so')

QS_OK = 0
QS_E_ARG = -1
QS_E_CHI = -2
QS_E_HIP = -3
QS_E_NOCONV = -4
QS_E_NODEV = -5
QS_E_RANGE = -6

QS_GATE_U1 = 0
QS_GATE_U2 = 1
QS_GATE_DIAG1 = 2
QS_GATE_DIAG2 = 3

QS_PREC_FP64 = 0
QS_PREC_FP32_MFMA = 1

QS_MPS_STATIC_SCRATCH_BYTES = 32 << 20

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)


class QsGate(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('q', ctypes.c_int32 * 2),
                ('reserved', ctypes.c_int32), ('ctrl_mask', ctypes.c_uint64),
                ('m', ctypes.c_double * 32)]


class QsMpsTask(ctypes.Structure):
    _fields_ = [('sites', ctypes.c_void_p), ('bond', ctypes.c_void_p),
                ('dlog', ctypes.c_void_p), ('site_stride', ctypes.c_int64),
                ('site', ctypes.c_int32), ('chi_max', ctypes.c_int32),
                ('cutoff', ctypes.c_double), ('n_sites', ctypes.c_int32),
                ('reserved', ctypes.c_int32), ('gate', ctypes.c_double * 32)]


class QsMpsBwdTask(ctypes.Structure):
    _fields_ = [('a', ctypes.c_void_p), ('b', ctypes.c_void_p), ('a2', ctypes.c_void_p),
                ('b2', ctypes.c_void_p), ('ga2', ctypes.c_void_p), ('gb2', ctypes.c_void_p),
                ('ga', ctypes.c_void_p), ('gb', ctypes.c_void_p), ('cl', ctypes.c_int32),
                ('cm', ctypes.c_int32), ('cr', ctypes.c_int32), ('k', ctypes.c_int32),
                ('eps', ctypes.c_double), ('gate', ctypes.c_double * 32)]


# name -> (restype, argtypes); every function in include/qsim.h
_PROTOS = {
    'qs_abi_version': (ctypes.c_int, []),
    'qs_init': (ctypes.c_int, [ctypes.c_int]),
    'qs_last_error': (ctypes.c_char_p, []),
    'qs_sv_max_qubits': (ctypes.c_int, []),
    'qs_sv_set_zero': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    'qs_sv_apply': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(QsGate),
                                   ctypes.c_int, ctypes.c_void_p]),
    'qs_sv_apply_words': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int32_p,
                                         ctypes.c_int, ctypes.c_void_p]),
    'qs_sv_plan_create': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(QsGate), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    'qs_sv_plan_launch': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    'qs_sv_plan_info': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_size_t)]),
    'qs_sv_plan_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'qs_sv_plan_specialize': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    'qs_sv_jit_source': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(QsGate), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]),
    'qs_sv_jit_check': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(QsGate), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_size_t)]),
    'qs_sv_plan_passes': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(QsGate), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]),
    'qs_sv_plan_json': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(QsGate), ctypes.c_int,
                                       ctypes.c_char_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    'qs_sv_time_passes': (ctypes.c_int, [ctypes.c_int]),
    'qs_sv_pass_times': (ctypes.c_int, [_c_double_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int)]),
    'qs_sv_norm_sq': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_double_p,
                                     ctypes.c_void_p]),
    'qs_sv_scale': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_void_p]),
    'qs_sv_reset_qubit': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p]),
    'qs_sv_probabilities': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    'qs_sv_sample_workspace_bytes': (ctypes.c_size_t, [ctypes.c_int]),
    'qs_sv_sample': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.c_void_p]),
    'qs_sv_pauli_inner': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         _c_int32_p, _c_int32_p, ctypes.c_int, _c_double_p,
                                         ctypes.c_void_p]),
    'qs_sv_inner': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_double_p,
                                   ctypes.c_void_p]),
    'qs_sv_apply_tfim': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_void_p]),
    'qs_mps_max_chi': (ctypes.c_int, []),
    'qs_mps_time_kernels': (ctypes.c_int, [ctypes.c_int]),
    'qs_mps_kernel_times': (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32), _c_double_p,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    'qs_mps_env_bytes': (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    'qs_mps_right_envs': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p]),
    'qs_mps_left_envs': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]),
    'qs_mps_entropy_work_bytes': (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    'qs_mps_entanglement_entropy': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_void_p, ctypes.c_void_p,
                                                   _c_int32_p, ctypes.c_int, ctypes.c_void_p,
                                                   ctypes.c_void_p, ctypes.c_size_t,
                                                   ctypes.c_void_p]),
    'qs_mps_expect_work_bytes': (ctypes.c_size_t, [ctypes.c_int]),
    'qs_mps_expect': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, _c_double_p, _c_double_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p]),
    'qs_mps_overlap_work_bytes': (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    'qs_mps_overlap': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_size_t, _c_double_p, ctypes.c_void_p]),
    'qs_mps_statevector': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_void_p]),
    'qs_mps_sample': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    'qs_mps_batch_table_bytes': (ctypes.c_size_t, [ctypes.c_int]),
    'qs_mps_batch_plan': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int)]),
    'qs_mps_batch_launch': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_void_p, ctypes.c_void_p]),
    'qs_mps_svd_lane_choices': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                               ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    'qs_mps_svd_lanes': (ctypes.c_int, [ctypes.c_int]),
    'qs_mps_set_svd_lanes': (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    'qs_mps_workspace_bytes': (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    'qs_mps_update_batch': (ctypes.c_int, [ctypes.POINTER(QsMpsTask), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_void_p, ctypes.c_void_p]),
    'qs_mps_update_batch_static': (ctypes.c_int, [ctypes.POINTER(QsMpsTask), ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    'qs_mps_apply_1q_batch': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_int, _c_int32_p, _c_double_p,
                                             ctypes.c_int, ctypes.c_void_p]),
    'qs_mps_reset_site': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    'qs_mps_backward_work_bytes': (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    'qs_mps_backward_batch': (ctypes.c_int, [ctypes.POINTER(QsMpsBwdTask), ctypes.c_int,
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    'qs_mps_apply_1q': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_int, _c_double_p, ctypes.c_void_p]),
}

_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None
_initialised_devices = set()


class QsError(RuntimeError):
    """A libqsim_hip call returned a non-zero status."""

    def __init__(self, func: str, status: int, message: str):
        super().__init__(f'{func} failed (status {status}): {message}')
        self.status = status


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libqsim_hip.so and declare every prototype (no device needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f'libqsim_hip.so not found at {path}: build it first '
                '(python -m qsim_amd.build, or __graft_entry__.build()). '
                'There is no CPU fallback.')
        lib = ctypes.CDLL(path)
        for name, (res, args) in _PROTOS.items():
            try:
                fn = getattr(lib, name)
            except AttributeError:
                continue            # reported by missing_symbols()
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def missing_symbols() -> list:
    """Prototypes of include/qsim.h that the loaded library does not export."""
    lb = lib()
    return [n for n in _PROTOS if not hasattr(lb, n)]


def lib() -> ctypes.CDLL:
    return _lib if _lib is not None else load_library()


def check(func: str, status: int) -> None:
    if status != QS_OK:
        msg = lib().qs_last_error()
        raise QsError(func, status, msg.decode() if msg else '')


def call(func: str, *args) -> int:
    st = getattr(lib(), func)(*args)
    check(func, st)
    return st


def ensure_device(device: int = 0) -> None:
    """qs_init(device) once per device; raises RuntimeError without a gfx950 GPU."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('no ROCm GPU visible: the HIP engine has no CPU fallback')
    if device in _initialised_devices:
        return
    call('qs_init', device)
    _initialised_devices.add(device)


def current_stream(device=None) -> ctypes.c_void_p:
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
