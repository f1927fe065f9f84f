"""ctypes front of the oracle's C/OpenMP statevector pair updates (sv_kernel_omp.c).

Test infrastructure and CPU baseline only (tests/ and bench.py's cpu_baseline leg).
The state is held as the FPGA kernel holds it: separate real and imaginary float64
arrays (opt_kernel.cpp:200-206).  Pinned by tests/test_oracle_sv.py against the NumPy
restatement (sv_ref.apply_kernel_ops).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import sv_ref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'libsv_kernel_omp.so')
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build_oracle import build
            build()
        _lib = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.sv_omp_apply_1q.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, dp]
        _lib.sv_omp_apply_cx.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib.sv_omp_apply_cz.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib.sv_omp_threads.restype = ctypes.c_int
    return _lib


def threads() -> int:
    return lib().sv_omp_threads()


def apply_kernel_ops(re: np.ndarray, im: np.ndarray, n: int, ops) -> None:
    """sv_ref.apply_kernel_ops on split arrays with the OpenMP pair updates."""
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    pr, pi = re.ctypes.data_as(dp), im.ctypes.data_as(dp)
    for g in ops:
        name = g[0].lower()
        qs = [int(v) for v in g[1]]
        ps = g[2] if len(g) > 2 else None
        op = sv_ref.KERNEL_GATE_OPS[name]
        if op == 17:
            L.sv_omp_apply_cx(pr, pi, n, qs[0], qs[1])
        elif op == 18:
            L.sv_omp_apply_cz(pr, pi, n, qs[0], qs[1])
        elif op != 0:
            pv = [float(v) for v in (ps or [])][:3] + [0.0, 0.0, 0.0]
            m = sv_ref.kernel_matrix(op, pv[0], pv[1], pv[2]).reshape(4)
            mv = np.ascontiguousarray(np.column_stack([m.real, m.imag]).reshape(8))
            L.sv_omp_apply_1q(pr, pi, n, qs[0], mv.ctypes.data_as(dp))
