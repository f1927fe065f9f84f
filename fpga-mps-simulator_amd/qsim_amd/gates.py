"""Gate tables of the two drop-in paths.

Two conventions exist in the reference and both are kept:

* MPS path (mps_core.py:24-144): named gates are 2x2 / 4x4 complex128 matrices;
  for a 4x4 gate on (q_first, q_second) the row index is
  2*bit(q_first) + bit(q_second).
* Statevector path (fpga_simulator_v07.py:243-325): gates are lowered to
  {17 one-qubit ops, cx, cz}.  The lowering is exact for every gate except
    - cp(theta): p, p, rzz(-theta/2) equals exp(i*theta/4) * CP(theta),
    - dcx(a, b): cx(a, b) then cx(b, a), which is the transpose of the
      MPS table's dcx.
  `sv_gate_records` produces direct gate records (no lowering) and returns the
  global phase the lowering would have produced, so amplitudes agree with the
  reference's output.
"""

from __future__ import annotations

import cmath
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import QS_GATE_DIAG1, QS_GATE_DIAG2, QS_GATE_U1, QS_GATE_U2

SQ2 = 1.0 / math.sqrt(2.0)

ONE_QUBIT_GATES = frozenset({'id', 'h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx', 'sxdg',
                             'rx', 'ry', 'rz', 'p', 'phase', 'u1', 'u2', 'u3', 'u'})
TWO_QUBIT_GATES = frozenset({'cx', 'cnot', 'cy', 'cz', 'ch', 'cp', 'cphase', 'crx', 'cry',
                             'crz', 'csx', 'swap', 'iswap', 'dcx', 'ecr', 'rxx', 'ryy', 'rzz'})


def _m(rows) -> np.ndarray:
    return np.asarray(rows, dtype=np.complex128)


_FIXED_1Q: Dict[str, np.ndarray] = {
    'id': _m([[1, 0], [0, 1]]),
    'h': _m([[SQ2, SQ2], [SQ2, -SQ2]]),
    'x': _m([[0, 1], [1, 0]]),
    'y': _m([[0, -1j], [1j, 0]]),
    'z': _m([[1, 0], [0, -1]]),
    's': _m([[1, 0], [0, 1j]]),
    'sdg': _m([[1, 0], [0, -1j]]),
    't': _m([[1, 0], [0, cmath.exp(1j * math.pi / 4)]]),
    'tdg': _m([[1, 0], [0, cmath.exp(-1j * math.pi / 4)]]),
    'sx': _m([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]),
    'sxdg': _m([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]]),
}

_DIAG_1Q = frozenset({'id', 'z', 's', 'sdg', 't', 'tdg', 'rz', 'p', 'phase', 'u1'})


def _params(params) -> List[float]:
    return [float(v) for v in (params or [])]


def matrix_1q(name: str, params: Optional[Sequence[float]] = None) -> np.ndarray:
    """2x2 of a one-qubit gate (mps_core.py:46-77; opt_kernel.cpp:257-347)."""
    nm = name.lower()
    if nm in _FIXED_1Q:
        return _FIXED_1Q[nm]
    p = _params(params)
    if nm in ('rx', 'ry', 'rz'):
        c, s = math.cos(0.5 * p[0]), math.sin(0.5 * p[0])
        if nm == 'rx':
            return _m([[c, -1j * s], [-1j * s, c]])
        if nm == 'ry':
            return _m([[c, -s], [s, c]])
        return _m([[complex(c, -s), 0], [0, complex(c, s)]])
    if nm in ('p', 'phase', 'u1'):
        return _m([[1, 0], [0, cmath.exp(1j * p[0])]])
    if nm == 'u2':
        phi, lam = p[0], p[1]
        return SQ2 * _m([[1, -cmath.exp(1j * lam)],
                         [cmath.exp(1j * phi), cmath.exp(1j * (phi + lam))]])
    if nm in ('u3', 'u'):
        th, phi, lam = p[0], p[1], p[2]
        c, s = math.cos(0.5 * th), math.sin(0.5 * th)
        return _m([[c, -cmath.exp(1j * lam) * s],
                   [cmath.exp(1j * phi) * s, cmath.exp(1j * (phi + lam)) * c]])
    raise ValueError(f"Unknown 1q gate: {name}")


def _ctrl(u: np.ndarray) -> np.ndarray:
    m = np.eye(4, dtype=np.complex128)
    m[2:, 2:] = u
    return m


def matrix_2q(name: str, params: Optional[Sequence[float]] = None) -> np.ndarray:
    """4x4 of a two-qubit gate in the MPS convention (mps_core.py:80-144)."""
    nm = name.lower()
    p = _params(params)
    if nm == 'cx':
        return _ctrl(_FIXED_1Q['x'])
    if nm == 'cy':
        return _ctrl(_FIXED_1Q['y'])
    if nm == 'cz':
        return _ctrl(_FIXED_1Q['z'])
    if nm == 'ch':
        return _ctrl(_FIXED_1Q['h'])
    if nm == 'csx':
        return _ctrl(_FIXED_1Q['sx'])
    if nm in ('crx', 'cry', 'crz'):
        return _ctrl(matrix_1q(nm[1:], p))
    if nm in ('cp', 'cphase'):
        return np.diag([1, 1, 1, cmath.exp(1j * p[0])]).astype(np.complex128)
    if nm == 'swap':
        return _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])
    if nm == 'iswap':
        return _m([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])
    if nm == 'dcx':
        return _m([[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]])
    if nm == 'ecr':
        return SQ2 * _m([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]])
    if nm in ('rxx', 'ryy'):
        c, s = math.cos(0.5 * p[0]), math.sin(0.5 * p[0])
        k = -1j * s if nm == 'rxx' else 1j * s
        return _m([[c, 0, 0, k], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [k, 0, 0, c]])
    if nm == 'rzz':
        e = cmath.exp(0.5j * p[0])
        return np.diag([e.conjugate(), e, e, e.conjugate()]).astype(np.complex128)
    raise ValueError(f"Unknown 2q gate: {name}")


def swap_qubit_order(g: np.ndarray) -> np.ndarray:
    """The same 4x4 gate written for the qubits in the other order."""
    perm = [0, 2, 1, 3]
    return g[np.ix_(perm, perm)]


# --------------------------------------------------------------------------
# statevector gate records
# --------------------------------------------------------------------------

def _record(kind: int, q0: int, q1: int = 0, ctrl: Sequence[int] = (),
            mat: Optional[np.ndarray] = None, diag: Optional[Sequence[complex]] = None):
    mask = 0
    for c in ctrl:
        mask |= 1 << int(c)
    vals = np.zeros(32)
    if mat is not None:
        flat = np.asarray(mat, dtype=np.complex128).reshape(-1)
        vals[0:2 * flat.size:2] = flat.real
        vals[1:2 * flat.size:2] = flat.imag
    if diag is not None:
        d = np.asarray(diag, dtype=np.complex128)
        vals[0:2 * d.size:2] = d.real
        vals[1:2 * d.size:2] = d.imag
    return (kind, int(q0), int(q1), mask, vals)


# SV path dcx = cx(a, b) followed by cx(b, a) (fpga_simulator_v07.py:282-284).
_DCX_SV = _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [0, 1, 0, 0]])


def sv_gate_records(name: str, qubits: Sequence[int], params=None,
                    nq: Optional[int] = None) -> Tuple[list, complex]:
    """(records, phase) for one logical gate of the statevector path.

    records: tuples (kind, q0, q1, ctrl_mask, m[32]) for qs_gate.
    phase:   global phase the reference lowering attaches to this gate.
    """
    nm = name.lower()
    qs = [int(q) for q in qubits]
    if nq is not None:
        for q in qs:
            if not 0 <= q < nq:
                raise ValueError(f"qubit {q} out of range for {nq} qubits")
    if len(set(qs)) != len(qs):
        raise ValueError(f"gate '{name}' repeats a qubit: {qs}")
    p = _params(params)
    if nm == 'id':
        return [], 1.0
    if nm in ONE_QUBIT_GATES:
        m = matrix_1q(nm, p)
        if nm in _DIAG_1Q:
            return [_record(QS_GATE_DIAG1, qs[0], diag=(m[0, 0], m[1, 1]))], 1.0
        return [_record(QS_GATE_U1, qs[0], mat=m)], 1.0
    if nm in ('cx', 'cnot'):
        return [_record(QS_GATE_U1, qs[1], ctrl=[qs[0]], mat=_FIXED_1Q['x'])], 1.0
    if nm == 'cy':
        return [_record(QS_GATE_U1, qs[1], ctrl=[qs[0]], mat=_FIXED_1Q['y'])], 1.0
    if nm == 'ch':
        return [_record(QS_GATE_U1, qs[1], ctrl=[qs[0]], mat=_FIXED_1Q['h'])], 1.0
    if nm == 'csx':
        return [_record(QS_GATE_U1, qs[1], ctrl=[qs[0]], mat=_FIXED_1Q['sx'])], 1.0
    if nm in ('crx', 'cry'):
        return [_record(QS_GATE_U1, qs[1], ctrl=[qs[0]], mat=matrix_1q(nm[1:], p))], 1.0
    if nm == 'crz':
        m = matrix_1q('rz', p)
        return [_record(QS_GATE_DIAG1, qs[1], ctrl=[qs[0]], diag=(m[0, 0], m[1, 1]))], 1.0
    if nm == 'cz':
        return [_record(QS_GATE_DIAG2, qs[0], qs[1], diag=(1, 1, 1, -1))], 1.0
    if nm in ('cp', 'cphase'):
        th = p[0]
        return ([_record(QS_GATE_DIAG2, qs[0], qs[1], diag=(1, 1, 1, cmath.exp(1j * th)))],
                cmath.exp(0.25j * th))
    if nm == 'rzz':
        m = matrix_2q('rzz', p)
        return [_record(QS_GATE_DIAG2, qs[0], qs[1], diag=np.diag(m))], 1.0
    if nm == 'dcx':
        return [_record(QS_GATE_U2, qs[0], qs[1], mat=_DCX_SV)], 1.0
    if nm in ('swap', 'iswap', 'ecr', 'rxx', 'ryy'):
        return [_record(QS_GATE_U2, qs[0], qs[1], mat=matrix_2q(nm, p))], 1.0
    if nm in ('ccx', 'toffoli', 'ccnot'):
        return [_record(QS_GATE_U1, qs[2], ctrl=qs[:2], mat=_FIXED_1Q['x'])], 1.0
    if nm in ('mcx', 'mct'):
        if not qs:
            raise ValueError("mcx needs at least one qubit")
        return [_record(QS_GATE_U1, qs[-1], ctrl=qs[:-1], mat=_FIXED_1Q['x'])], 1.0
    raise ValueError(f"Unsupported gate '{name}'")
