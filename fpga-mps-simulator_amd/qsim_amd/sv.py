"""Statevector engine over libqsim_hip.so.

The state lives in HBM as one complex128 torch tensor of 2^nq amplitudes
(index bit q = qubit q).  Gates go to the library as qs_gate records and are
applied by fused passes (one HBM read + write of the state per pass).
"""

from __future__ import annotations

import ctypes
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .gates import sv_gate_records


def records_to_array(records: Sequence[tuple]):
    arr = (_lib.QsGate * max(len(records), 1))()
    for i, (kind, q0, q1, mask, vals) in enumerate(records):
        g = arr[i]
        g.kind = kind
        g.q[0] = q0
        g.q[1] = q1
        g.ctrl_mask = mask
        g.m[:] = [float(v) for v in vals]
    return arr


def gates_to_records(gates: Iterable[tuple], nq: int) -> Tuple[list, complex]:
    """Logical (name, qubits, params) tuples -> qs_gate records + total phase."""
    recs: list = []
    phase = 1.0 + 0.0j
    for g in gates:
        name, qs = g[0], g[1]
        ps = g[2] if len(g) > 2 else None
        r, ph = sv_gate_records(name, qs, ps, nq)
        recs.extend(r)
        phase *= ph
    return recs, phase


class SvPlan:
    """A fused-pass plan built once on the host and kept in device memory
    (qs_sv_plan_create); StateVector.launch applies it with no host planning or upload."""

    def __init__(self, nq: int, records: Sequence[tuple], device: int = 0):
        _lib.ensure_device(device)
        arr = records_to_array(records)
        h = ctypes.c_void_p()
        _lib.call('qs_sv_plan_create', nq, arr, len(records), ctypes.byref(h))
        self._h = h
        self.nq = nq
        self.device = device
        n, p, b = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_size_t(0)
        _lib.call('qs_sv_plan_info', h, ctypes.byref(n), ctypes.byref(p), ctypes.byref(b))
        self.npasses = p.value
        self.device_bytes = b.value

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError('plan already destroyed')
        return self._h

    def specialize(self) -> int:
        """Compile this plan's per-pass kernels (qs_sv_plan_specialize, hipRTC); later
        launches run them.  Returns the number of kernels compiled by this call (0 when
        the process-wide cache already held them all)."""
        n = ctypes.c_int(0)
        _lib.call('qs_sv_plan_specialize', self.handle, ctypes.byref(n))
        return n.value

    def destroy(self) -> None:
        if getattr(self, '_h', None) is not None:
            _lib.call('qs_sv_plan_destroy', self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class StateVector:
    """A 2^nq complex128 state on one GPU."""

    def __init__(self, nq: int, device: int = 0, init_zero: bool = True):
        import torch
        _lib.ensure_device(device)
        maxq = _lib.lib().qs_sv_max_qubits()
        if not 1 <= nq <= maxq:
            raise ValueError(f"num_qubits must be in [1,{maxq}]")
        self.nq = nq
        self.device = device
        self.tensor = torch.empty(1 << nq, dtype=torch.complex128, device=f'cuda:{device}')
        if init_zero:
            self.set_zero()

    # -- plumbing ---------------------------------------------------------
    @property
    def ptr(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.tensor.data_ptr())

    def _stream(self):
        return _lib.current_stream(self.device)

    # -- state management -----------------------------------------------------
    def set_zero(self) -> None:
        _lib.call('qs_sv_set_zero', self.ptr, self.nq, self._stream())

    def load(self, sv: np.ndarray) -> None:
        import torch
        arr = np.ascontiguousarray(np.asarray(sv, dtype=np.complex128).reshape(-1))
        if arr.size != (1 << self.nq):
            raise ValueError(f"state has {arr.size} amplitudes, expected {1 << self.nq}")
        self.tensor.copy_(torch.from_numpy(arr))

    def to_numpy(self) -> np.ndarray:
        return self.tensor.cpu().numpy()

    # -- gates -------------------------------------------------------------
    def apply_records(self, records: Sequence[tuple]) -> None:
        if not records:
            return
        arr = records_to_array(records)
        _lib.call('qs_sv_apply', self.ptr, self.nq, arr, len(records), self._stream())

    def plan(self, records: Sequence[tuple]) -> SvPlan:
        """Plan `records` once for this state's qubit count (see SvPlan)."""
        return SvPlan(self.nq, records, self.device)

    def launch(self, plan: SvPlan) -> None:
        """Apply a prebuilt plan on the current stream (no planning, no upload)."""
        if plan.nq != self.nq:
            raise ValueError(f'plan is for {plan.nq} qubits, state has {self.nq}')
        _lib.call('qs_sv_plan_launch', plan.handle, self.ptr, self._stream())

    def apply(self, gates: Iterable[tuple]) -> complex:
        """Apply logical gates; returns the reference-lowering global phase."""
        recs, phase = gates_to_records(gates, self.nq)
        self.apply_records(recs)
        return phase

    def apply_words(self, words: np.ndarray) -> None:
        """Apply gates in the reference's 8-word instruction format."""
        w = np.ascontiguousarray(np.asarray(words, dtype=np.int32).reshape(-1))
        if w.size % 8:
            raise ValueError("gate_sequence length must be a multiple of 8")
        _lib.call('qs_sv_apply_words', self.ptr, self.nq,
                  w.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), w.size // 8,
                  self._stream())

    def norm_sq(self) -> float:
        out = ctypes.c_double(0.0)
        _lib.call('qs_sv_norm_sq', self.ptr, self.nq, ctypes.byref(out), self._stream())
        return out.value

    def scale(self, factor: complex) -> None:
        f = complex(factor)
        _lib.call('qs_sv_scale', self.ptr, self.nq, f.real, f.imag, self._stream())

    def reset_qubit(self, q: int) -> None:
        _lib.call('qs_sv_reset_qubit', self.ptr, self.nq, int(q), self._stream())

    def probabilities(self):
        import torch
        p = torch.empty(1 << self.nq, dtype=torch.float64, device=self.tensor.device)
        _lib.call('qs_sv_probabilities', self.ptr, self.nq, ctypes.c_void_p(p.data_ptr()),
                  self._stream())
        return p

    def sample_indices(self, uniforms) -> np.ndarray:
        """Basis-state indices for the given uniforms in [0, 1): for each, the first index
        whose cumulative probability |a|^2 / sum |a|^2 exceeds it, which is what
        np.random.choice(2^n, len(uniforms), p=probs) returns for the uniforms it draws
        (fpga_simulator_v07.py:592-609).  Computed on the device."""
        import torch
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64).reshape(-1))
        if u.size == 0:
            return np.zeros(0, dtype=np.int64)
        if not (np.all(u >= 0.0) and np.all(u < 1.0)):
            raise ValueError('uniforms must lie in [0, 1)')
        dev = self.tensor.device
        ud = torch.from_numpy(u).to(dev)
        out = torch.empty(u.size, dtype=torch.int64, device=dev)
        nbytes = _lib.lib().qs_sv_sample_workspace_bytes(self.nq)
        ws = torch.empty(max(1, (nbytes + 7) // 8), dtype=torch.float64, device=dev)
        _lib.call('qs_sv_sample', self.ptr, self.nq, ctypes.c_void_p(ud.data_ptr()), u.size,
                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                  ws.numel() * 8, self._stream())
        return out.cpu().numpy()


def time_passes(enable: bool) -> None:
    """Bracket every fused-pass kernel launched from this thread with HIP events."""
    _lib.call('qs_sv_time_passes', 1 if enable else 0)


def pass_times(cap: int = 1 << 16) -> List[float]:
    """Durations (ms) of the pass kernels recorded since the last call; waits for them."""
    buf = (ctypes.c_double * cap)()
    n = ctypes.c_int(0)
    _lib.call('qs_sv_pass_times', buf, cap, ctypes.byref(n))
    return [buf[i] for i in range(min(n.value, cap))]


def plan_passes(nq: int, records: Sequence[tuple]) -> int:
    arr = records_to_array(records)
    n = ctypes.c_int(0)
    _lib.call('qs_sv_plan_passes', nq, arr, len(records), ctypes.byref(n))
    return n.value


def jit_check(nq: int, records: Sequence[tuple]) -> Tuple[int, int]:
    """Generate and compile the per-pass kernels of the plan of `records` without a
    device (qs_sv_jit_check): (distinct kernels, bytes of code objects)."""
    arr = records_to_array(records)
    nk, nb = ctypes.c_int(0), ctypes.c_size_t(0)
    _lib.call('qs_sv_jit_check', nq, arr, len(records), ctypes.byref(nk), ctypes.byref(nb))
    return nk.value, nb.value


def jit_source(nq: int, records: Sequence[tuple], pass_index: int) -> str:
    """Source text of one pass's specialised kernel (qs_sv_jit_source, host only)."""
    arr = records_to_array(records)
    need = ctypes.c_size_t(0)
    _lib.call('qs_sv_jit_source', nq, arr, len(records), pass_index, None, 0,
              ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    _lib.call('qs_sv_jit_source', nq, arr, len(records), pass_index, buf, need.value,
              ctypes.byref(need))
    return buf.value.decode()


def plan_json(nq: int, records: Sequence[tuple]) -> dict:
    """The fused-pass plan (host only; usable without a GPU)."""
    import json
    arr = records_to_array(records)
    need = ctypes.c_size_t(0)
    _lib.call('qs_sv_plan_json', nq, arr, len(records), None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    _lib.call('qs_sv_plan_json', nq, arr, len(records), buf, need.value, ctypes.byref(need))
    return json.loads(buf.value.decode())
