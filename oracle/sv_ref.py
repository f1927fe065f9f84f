"""CPU oracle for the statevector gate sweep (test infrastructure only).

Imported by `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg
of `bench.py`; never by the product package.  It restates in NumPy FP64:

  * the SV kernel's opcode table and per-opcode 2x2 matrices
    (opt_kernel.cpp:39-58, 257-347),
  * the three pair updates: dense one-qubit pass, CX swap pass, CZ negate
    pass (opt_kernel.cpp:71-195),
  * the host lowering to {17 one-qubit ops, cx, cz}
    (fpga_simulator_v07.py:243-325) and the segment splitter
    (fpga_simulator_v07.py:330-369),
  * run_circuit's state handling: |0> init, host reset projection with the
    |0> fallback, final normalisation, clbit-mapped sampling
    (fpga_simulator_v07.py:502-513, 533-612).

Deviation kept on purpose: the FPGA kernel receives parameters as float32
(fpga_simulator_v07.py:469-472, opt_kernel.cpp:248-255).  `run_circuit`
takes `f32_params`; the default False evaluates them in FP64, which is what
the HIP path does.

mcx with more than two controls is lowered by Qiskit in the reference
(fpga_simulator_v07.py:205-231), which is not installed here; the oracle
applies the exact multi-controlled X permutation instead and is pinned by
the expected indices of mcx_stress_results.csv.

Pinning: every kernel matrix is checked against closed forms and the MPS
oracle's gate table; the pair updates against dense Kronecker products at
n <= 6; the lowering against the exact gate matrices up to the documented
global phase; run_circuit against the comprehensive_test_v07.py KATs.
"""

from __future__ import annotations

import math
from collections import defaultdict
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

KERNEL_GATE_OPS: Dict[str, int] = {
    'id': 0, 'h': 1, 'x': 2, 'y': 3, 'z': 4, 's': 5, 't': 6, 'sdg': 7, 'tdg': 8,
    'sx': 9, 'sxdg': 10, 'rx': 11, 'ry': 12, 'rz': 13, 'p': 14, 'phase': 14, 'u1': 14,
    'u2': 15, 'u3': 16, 'u': 16, 'cx': 17, 'cnot': 17, 'cz': 18,
}                                                   # fpga_simulator_v07.py:69-78

ALL_SUPPORTED_GATES = frozenset({
    'id', 'h', 'x', 'y', 'z', 's', 't', 'sdg', 'tdg', 'sx', 'sxdg',
    'rx', 'ry', 'rz', 'p', 'phase', 'u1', 'u2', 'u3', 'u',
    'cx', 'cnot', 'cz', 'cy', 'ch', 'swap', 'cp', 'cphase',
    'crx', 'cry', 'crz', 'csx', 'dcx', 'iswap', 'ecr', 'rxx', 'ryy', 'rzz',
    'ccx', 'toffoli', 'ccnot', 'mcx', 'mct',
})                                                  # fpga_simulator_v07.py:84-93
NON_UNITARY = frozenset({'measure', 'barrier', 'snapshot', 'delay', 'initialize'})

_R2 = 0.7071067811865476


def kernel_matrix(op: int, p0: float = 0.0, p1: float = 0.0, p2: float = 0.0) -> np.ndarray:
    """2x2 matrix the SV kernel builds for a one-qubit opcode (opt_kernel.cpp:257-347)."""
    if op == 0:
        return np.eye(2, dtype=np.complex128)
    if op == 1:
        return np.array([[_R2, _R2], [_R2, -_R2]], dtype=np.complex128)
    if op == 2:
        return np.array([[0, 1], [1, 0]], dtype=np.complex128)
    if op == 3:
        return np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    if op == 4:
        return np.array([[1, 0], [0, -1]], dtype=np.complex128)
    if op == 5:
        return np.array([[1, 0], [0, 1j]], dtype=np.complex128)
    if op == 6:
        return np.array([[1, 0], [0, _R2 + 1j * _R2]], dtype=np.complex128)
    if op == 7:
        return np.array([[1, 0], [0, -1j]], dtype=np.complex128)
    if op == 8:
        return np.array([[1, 0], [0, _R2 - 1j * _R2]], dtype=np.complex128)
    if op == 9:
        return np.array([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])
    if op == 10:
        return np.array([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])
    if op in (11, 12, 13):
        c, s = math.cos(0.5 * p0), math.sin(0.5 * p0)
        if op == 11:
            return np.array([[c, -1j * s], [-1j * s, c]])
        if op == 12:
            return np.array([[c, -s], [s, c]], dtype=np.complex128)
        return np.array([[c - 1j * s, 0], [0, c + 1j * s]])
    if op == 14:
        return np.array([[1, 0], [0, complex(math.cos(p0), math.sin(p0))]])
    if op == 15:
        return _R2 * np.array([[1, -complex(math.cos(p1), math.sin(p1))],
                               [complex(math.cos(p0), math.sin(p0)),
                                complex(math.cos(p0 + p1), math.sin(p0 + p1))]])
    if op == 16:
        ct, st = math.cos(0.5 * p0), math.sin(0.5 * p0)
        return np.array([[ct, -st * complex(math.cos(p2), math.sin(p2))],
                         [st * complex(math.cos(p1), math.sin(p1)),
                          ct * complex(math.cos(p1 + p2), math.sin(p1 + p2))]])
    raise ValueError(f"not a one-qubit opcode: {op}")


# ----------------------------------------------------------------------------
# Pair updates (opt_kernel.cpp:71-195).  Index bit q of the flat state is qubit q.
# ----------------------------------------------------------------------------

def apply_1q(state: np.ndarray, t: int, m: np.ndarray) -> None:
    """Dense one-qubit pass: (a[i0], a[i1]) <- m @ (a[i0], a[i1]), i1 = i0 + 2^t."""
    v = state.reshape(-1, 2, 1 << t)
    a0 = v[:, 0, :].copy()
    a1 = v[:, 1, :].copy()
    v[:, 0, :] = m[0, 0] * a0 + m[0, 1] * a1
    v[:, 1, :] = m[1, 0] * a0 + m[1, 1] * a1


def _bit_view(state: np.ndarray, n: int) -> np.ndarray:
    return state.reshape([2] * n)   # axis k is qubit n-1-k


def apply_cx(state: np.ndarray, n: int, c: int, t: int) -> None:
    """Swap the target pair wherever the control bit is 1."""
    v = _bit_view(state, n)
    sel = [slice(None)] * n
    sel[n - 1 - c] = 1
    sub = v[tuple(sel)]                              # control fixed to 1
    ax = (n - 1 - t) - (1 if (n - 1 - c) < (n - 1 - t) else 0)
    sub[...] = np.flip(sub, axis=ax).copy()


def apply_cz(state: np.ndarray, n: int, c: int, t: int) -> None:
    """Negate amplitudes with both bits set."""
    v = _bit_view(state, n)
    sel = [slice(None)] * n
    sel[n - 1 - c] = 1
    sel[n - 1 - t] = 1
    v[tuple(sel)] *= -1


def apply_mcx(state: np.ndarray, n: int, ctrls: Sequence[int], t: int) -> None:
    """Exact multi-controlled X (permutation), used for mcx with > 2 controls."""
    idx = np.arange(state.size, dtype=np.int64)
    mask = 0
    for c in ctrls:
        mask |= 1 << c
    sel = idx[((idx & mask) == mask) & (((idx >> t) & 1) == 0)]
    tmp = state[sel].copy()
    state[sel] = state[sel | (1 << t)]
    state[sel | (1 << t)] = tmp


# ----------------------------------------------------------------------------
# Lowering (fpga_simulator_v07.py:236-325) and splitting (:330-369)
# ----------------------------------------------------------------------------

_PASSTHROUGH = frozenset({'id', 'h', 'x', 'y', 'z', 's', 't', 'sdg', 'tdg', 'sx', 'sxdg',
                          'rx', 'ry', 'rz', 'p', 'phase', 'u1', 'u2', 'u3', 'u',
                          'cx', 'cnot', 'cz', 'measure', 'barrier', 'snapshot', 'delay',
                          'initialize', 'reset', 'mcx', 'mct'})


def lower(gates):
    """Rewrite gate tuples into {one-qubit natives, cx, cz}, rule by rule as the
    reference host lowering does (the rules themselves are part of the
    reference's observable semantics: cp picks up a global phase e^{i theta/4})."""
    out = []
    pi = math.pi
    for g in gates:
        name = g[0].lower() if isinstance(g[0], str) else g[0]
        q = g[1]
        p = g[2] if len(g) > 2 else None
        if name in _PASSTHROUGH:
            out.append(g)
            continue
        if name == 'u':
            out.append(('u3', q, p))
            continue
        if name in ('u1', 'phase'):
            out.append(('p', q, p))
            continue
        a = int(q[0])
        b = int(q[1]) if len(q) > 1 else None
        th = float(p[0]) if p else 0.0
        if name == 'swap':
            out += lower([('cx', [a, b], None), ('cx', [b, a], None), ('cx', [a, b], None)])
        elif name == 'cy':
            out += lower([('sdg', [b], None), ('cx', [a, b], None), ('s', [b], None)])
        elif name == 'ch':
            out += lower([('ry', [b], [-pi / 4]), ('cz', [a, b], None), ('ry', [b], [pi / 4])])
        elif name == 'rzz':
            out += lower([('cx', [a, b], None), ('rz', [b], [th]), ('cx', [a, b], None)])
        elif name in ('cp', 'cphase'):
            out.append(('p', [a], [th / 2]))
            out.append(('p', [b], [th / 2]))
            out += lower([('rzz', [a, b], [-th / 2])])
        elif name == 'crz':
            out += lower([('cx', [a, b], None), ('rz', [b], [-th / 2]),
                          ('cx', [a, b], None), ('rz', [b], [th / 2])])
        elif name == 'crx':
            out.append(('h', [b], None))
            out += lower([('crz', [a, b], [th])])
            out.append(('h', [b], None))
        elif name == 'cry':
            out += lower([('cx', [a, b], None), ('ry', [b], [-th / 2]),
                          ('cx', [a, b], None), ('ry', [b], [th / 2])])
        elif name == 'csx':
            out += lower([('crx', [a, b], [pi / 2])])
            out.append(('p', [a], [pi / 4]))
        elif name == 'dcx':
            out += lower([('cx', [a, b], None), ('cx', [b, a], None)])
        elif name == 'iswap':
            out += lower([('s', [a], None), ('s', [b], None), ('h', [a], None),
                          ('cx', [a, b], None), ('cx', [b, a], None), ('h', [b], None)])
        elif name == 'ecr':
            out.append(('h', [b], None))
            out += lower([('cx', [a, b], None), ('rz', [b], [pi / 4]), ('cx', [a, b], None)])
            out += [('h', [b], None), ('x', [a], None), ('h', [b], None)]
            out += lower([('cx', [a, b], None), ('rz', [b], [-pi / 4]), ('cx', [a, b], None)])
            out.append(('h', [b], None))
        elif name == 'rxx':
            out += [('h', [a], None), ('h', [b], None)]
            out += lower([('cx', [a, b], None), ('rz', [b], [th]), ('cx', [a, b], None)])
            out += [('h', [a], None), ('h', [b], None)]
        elif name == 'ryy':
            out += [('sdg', [a], None), ('sdg', [b], None), ('h', [a], None), ('h', [b], None)]
            out += lower([('cx', [a, b], None), ('rz', [b], [th]), ('cx', [a, b], None)])
            out += [('h', [a], None), ('h', [b], None), ('s', [a], None), ('s', [b], None)]
        elif name in ('ccx', 'toffoli', 'ccnot'):
            c0, c1, t = int(q[0]), int(q[1]), int(q[2])
            out += lower([('h', [t], None), ('cx', [c1, t], None), ('tdg', [t], None),
                          ('cx', [c0, t], None), ('t', [t], None), ('cx', [c1, t], None),
                          ('tdg', [t], None), ('cx', [c0, t], None), ('t', [c1], None),
                          ('t', [t], None), ('h', [t], None), ('cx', [c0, c1], None),
                          ('t', [c0], None), ('tdg', [c1], None), ('cx', [c0, c1], None)])
        else:
            out.append(g)
    return out


def split_segments(gates):
    """(segments, resets, meas) of the reference splitter; meas is (qubit, clbit)."""
    segments: List[list] = [[]]
    resets: List[Tuple[int, int]] = []
    meas: List[Tuple[int, int]] = []
    for g in gates:
        if len(g) == 2:
            name, qs = g
            ps = None
        elif len(g) == 3:
            name, qs, ps = g
        else:
            raise ValueError(f"Bad gate tuple: {g!r}")
        if not isinstance(name, str):
            raise ValueError("Gate name must be str")
        low = name.lower()
        if low in NON_UNITARY:
            if low == 'measure':
                cl = list(ps) if ps else []
                for k, qq in enumerate(list(qs or [])):
                    meas.append((int(qq), int(cl[k]) if k < len(cl) else int(qq)))
            continue
        if low == 'reset':
            for qq in (qs or []):
                resets.append((int(qq), len(segments) - 1))
            segments.append([])
            continue
        if low not in ALL_SUPPORTED_GATES:
            raise ValueError(f"Unsupported gate '{name}'")
        if low in ('mcx', 'mct'):
            qs = [int(v) for v in qs]
            if not qs:
                raise ValueError("mcx needs at least one qubit")
            t, ctrls = qs[-1], qs[:-1]
            if t in ctrls:
                raise ValueError(f"mcx target {t} in controls {ctrls}")
            if len(ctrls) == 0:
                segments[-1].append(('x', [t], None))
            elif len(ctrls) == 1:
                segments[-1].append(('cx', [ctrls[0], t], None))
            elif len(ctrls) == 2:
                segments[-1].append(('ccx', [ctrls[0], ctrls[1], t], None))
            else:
                segments[-1].append(('__mcx__', ctrls + [t], None))
        else:
            segments[-1].append((low, list(qs), ps))
    if segments and not segments[-1]:
        segments.pop()
    return segments, resets, meas


def apply_kernel_ops(state: np.ndarray, n: int, ops, f32_params: bool = False) -> None:
    """Apply lowered ops with the kernel's pass semantics."""
    for g in ops:
        name = g[0].lower()
        qs = [int(v) for v in g[1]]
        ps = g[2] if len(g) > 2 else None
        if name in ('measure', 'barrier', 'snapshot', 'delay', 'initialize'):
            continue
        if name == '__mcx__':
            apply_mcx(state, n, qs[:-1], qs[-1])
            continue
        op = KERNEL_GATE_OPS[name]
        if op == 17:
            apply_cx(state, n, qs[0], qs[1])
        elif op == 18:
            apply_cz(state, n, qs[0], qs[1])
        elif op == 0:
            continue
        else:
            pv = [float(v) for v in (ps or [])][:3] + [0.0, 0.0, 0.0]
            if f32_params:
                pv = [float(np.float32(v)) for v in pv]
            apply_1q(state, qs[0], kernel_matrix(op, pv[0], pv[1], pv[2]))


def run_circuit(nq: int, gates, f32_params: bool = False,
                init: Optional[np.ndarray] = None) -> np.ndarray:
    """Final normalised statevector of the reference run_circuit (no sampling)."""
    segments, resets, _ = split_segments(gates)
    segments = [lower(s) for s in segments]
    if init is None:
        state = np.zeros(1 << nq, dtype=np.complex128)
        state[0] = 1.0
    else:
        state = np.array(init, dtype=np.complex128).copy()
    by_seg = defaultdict(list)
    for q, si in resets:
        by_seg[si].append(q)
    for si, seg in enumerate(segments):
        apply_kernel_ops(state, nq, seg, f32_params)
        for q in by_seg.get(si, []):
            reset_qubit(state, q)
    nb = float(np.linalg.norm(state))
    if nb > 1e-12:
        state /= nb
    return state


def reset_qubit(state: np.ndarray, q: int) -> None:
    """Host reset projection with |0> fallback (fpga_simulator_v07.py:502-513)."""
    idx = np.arange(state.size, dtype=np.int64)
    state[(idx >> q) & 1 == 1] = 0.0
    nb = float(np.linalg.norm(state))
    if nb < 1e-12:
        state[:] = 0.0
        state[0] = 1.0
    else:
        state /= nb


def clbit_key(index: int, nq: int, nc: int, meas_pairs) -> str:
    """Count key for a sampled basis index (fpga_simulator_v07.py:602-608)."""
    qbits = format(int(index), f'0{nq}b')
    cb = ['0'] * nc
    for q, c in meas_pairs:
        if 0 <= q < nq and 0 <= c < nc:
            cb[nc - 1 - c] = qbits[nq - 1 - q]
    return ''.join(cb)


# ----------------------------------------------------------------------------
# Workload generators (inputs only)
# ----------------------------------------------------------------------------

def qft_gates(n: int, inverse: bool = False):
    """QFT gate list in the ordering of comprehensive_test_v07.py:433-452."""
    g = []
    if not inverse:
        for i in range(n):
            g.append(('h', [i], None))
            for j in range(i + 1, n):
                g.append(('cp', [j, i], [math.pi / (2 ** (j - i))]))
        for i in range(n // 2):
            g.append(('swap', [i, n - 1 - i], None))
    else:
        for i in range(n // 2):
            g.append(('swap', [i, n - 1 - i], None))
        for i in range(n - 1, -1, -1):
            for j in range(n - 1, i, -1):
                g.append(('cp', [j, i], [-math.pi / (2 ** (j - i))]))
            g.append(('h', [i], None))
    return g


def u3_cz_layers(n: int, depth: int, seed: int):
    """depth x [U3(theta,phi,lam ~ U[0,2pi)) on every qubit, then CZ(i,i+1) ladder]."""
    rng = np.random.default_rng(seed)
    g = []
    for _ in range(depth):
        ang = rng.uniform(0.0, 2 * math.pi, size=(n, 3))
        for q in range(n):
            g.append(('u3', [q], [float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2])]))
        for q in range(n - 1):
            g.append(('cz', [q, q + 1], None))
    return g


def sv_benchmark_circuit(n: int = 28, depth: int = 20, seed: int = 1234):
    """depth layers of U3 + CZ ladder followed by an n-qubit QFT."""
    return u3_cz_layers(n, depth, seed) + qft_gates(n)


# ----------------------------------------------------------------------------
# TFIM energy (builder-added observable; no reference counterpart)
# ----------------------------------------------------------------------------

def tfim_energy_sv(state: np.ndarray, j_coupling: float = 1.0, h_field: float = 1.0) -> float:
    """<psi|H|psi>/<psi|psi> for H = -J sum Z_i Z_{i+1} - h sum X_i (open chain)."""
    n = int(round(math.log2(state.size)))
    idx = np.arange(state.size, dtype=np.int64)
    p = np.abs(state) ** 2
    zz = np.zeros(state.size)
    for i in range(n - 1):
        zz += 1.0 - 2.0 * (((idx >> i) ^ (idx >> (i + 1))) & 1)
    e = -j_coupling * float(np.dot(p, zz))
    for i in range(n):
        e -= h_field * float(np.real(np.vdot(state, state[idx ^ (1 << i)])))
    return e / float(p.sum())


def hea_gates(n: int, theta, layers: int = 4):
    """layers x [Ry(theta) every qubit, Rz(theta) every qubit, CZ ladder]."""
    g = []
    k = 0
    for _ in range(layers):
        for q in range(n):
            g.append(('ry', [q], [float(theta[k + q])]))
        k += n
        for q in range(n):
            g.append(('rz', [q], [float(theta[k + q])]))
        k += n
        for q in range(n - 1):
            g.append(('cz', [q, q + 1], None))
    return g
