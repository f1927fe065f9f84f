"""Adjoint gradients of <psi(theta)|H|psi(theta)> on the statevector path.

A parameterised circuit is a list of blocks.  A rotation block holds rotations
exp(-i theta_k P_k / 2) (P in {X, Y, Z}: rx, ry, rz) on distinct qubits, so they
commute; a fixed block holds any gates without parameters.

Forward: psi = U(theta)|0>, E = <psi|H psi>.  Backward (reverse sweep): with
lam = H psi, walk the blocks backwards; at a rotation block every gradient is
dE/dtheta_k = Im <lam|P_k|psi> (one multi-Pauli reduction over both states,
qs_sv_pauli_inner), then the block is undone on psi and lam with its daggered
gates (fused passes).  Cost: one forward + two reverse sweeps + one reduction per
rotation block, independent of the number of parameters per block.
"""

from __future__ import annotations

import ctypes
import math
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib
from .sv import StateVector

_PAULI_OF = {'rx': 0, 'ry': 1, 'rz': 2}


class Block:
    def __init__(self, kind: str, ops: list):
        self.kind = kind          # 'rot' or 'fixed'
        self.ops = ops            # rot: [(name, qubit, param_index)]; fixed: gate tuples


class ParamCircuit:
    """Blocks of rotations and fixed gates on nq qubits."""

    def __init__(self, nq: int, blocks: List[Block], nparams: int):
        self.nq = nq
        self.blocks = blocks
        self.nparams = nparams
        for b in blocks:
            if b.kind == 'rot':
                qs = [q for _, q, _ in b.ops]
                if len(set(qs)) != len(qs):
                    raise ValueError('a rotation block must act on distinct qubits')
                for nm, _, _ in b.ops:
                    if nm not in _PAULI_OF:
                        raise ValueError(f'rotation {nm} is not rx/ry/rz')

    def gates(self, theta: Sequence[float], inverse_block=None) -> list:
        out = []
        for b in self.blocks:
            out += self.block_gates(b, theta)
        return out

    @staticmethod
    def block_gates(b: Block, theta, dagger: bool = False) -> list:
        if b.kind == 'rot':
            s = -1.0 if dagger else 1.0
            return [(nm, [q], [s * float(theta[k])]) for nm, q, k in b.ops]
        if not dagger:
            return list(b.ops)
        inv = []
        for g in reversed(b.ops):
            nm = g[0].lower()
            if nm in ('cz', 'cx', 'cnot', 'h', 'x', 'y', 'z', 'swap', 'ccx'):
                inv.append(g)
            elif nm in ('rx', 'ry', 'rz', 'p', 'cp', 'rzz', 'crz', 'crx', 'cry'):
                inv.append((nm, g[1], [-float(g[2][0])]))
            elif nm in ('s', 't'):
                inv.append((nm + 'dg', g[1], None))
            elif nm in ('sdg', 'tdg'):
                inv.append((nm[:-2], g[1], None))
            else:
                raise ValueError(f'no dagger rule for fixed gate {nm}')
        return inv


def hardware_efficient_ansatz(nq: int, layers: int = 4) -> ParamCircuit:
    """layers x [Ry on every qubit, Rz on every qubit, CZ(i, i+1) ladder]."""
    blocks = []
    k = 0
    for _ in range(layers):
        blocks.append(Block('rot', [('ry', q, k + q) for q in range(nq)]))
        k += nq
        blocks.append(Block('rot', [('rz', q, k + q) for q in range(nq)]))
        k += nq
        blocks.append(Block('fixed', [('cz', [q, q + 1], None) for q in range(nq - 1)]))
    return ParamCircuit(nq, blocks, k)


class TFIMEnergy:
    """E(theta) = <psi|H|psi>, H = -J sum Z_i Z_{i+1} - h sum X_i, and its adjoint gradient."""

    def __init__(self, circuit: ParamCircuit, J: float = 1.0, h: float = 1.0, device: int = 0):
        self.c = circuit
        self.J = float(J)
        self.h = float(h)
        self.device = device
        self.psi = StateVector(circuit.nq, device)
        self.lam = StateVector(circuit.nq, device, init_zero=False)

    def _stream(self):
        return _lib.current_stream(self.device)

    def _hpsi(self):
        _lib.call('qs_sv_apply_tfim', self.psi.ptr, self.lam.ptr, self.c.nq, self.J, self.h,
                  self._stream())

    def _inner(self, a: StateVector, b: StateVector) -> complex:
        out = (ctypes.c_double * 2)()
        _lib.call('qs_sv_inner', a.ptr, b.ptr, self.c.nq, out, self._stream())
        return complex(out[0], out[1])

    def energy(self, theta) -> float:
        self.psi.set_zero()
        self.psi.apply(self.c.gates(theta))
        self._hpsi()
        return self._inner(self.psi, self.lam).real

    def energy_and_grad(self, theta) -> Tuple[float, np.ndarray]:
        theta = np.asarray(theta, dtype=np.float64)
        e = self.energy(theta)
        grad = np.zeros(self.c.nparams)
        for b in reversed(self.c.blocks):
            if b.kind == 'rot':
                kinds = (ctypes.c_int32 * len(b.ops))(*[_PAULI_OF[nm] for nm, _, _ in b.ops])
                qubits = (ctypes.c_int32 * len(b.ops))(*[q for _, q, _ in b.ops])
                out = (ctypes.c_double * (2 * len(b.ops)))()
                _lib.call('qs_sv_pauli_inner', self.lam.ptr, self.psi.ptr, self.c.nq, kinds,
                          qubits, len(b.ops), out, self._stream())
                for j, (_, _, k) in enumerate(b.ops):
                    grad[k] = out[2 * j + 1]
            dag = ParamCircuit.block_gates(b, theta, dagger=True)
            self.psi.apply(dag)
            self.lam.apply(dag)
        return e, grad


def torch_energy(module: TFIMEnergy):
    """A torch.autograd.Function computing E(theta) with the adjoint backward."""
    import torch

    class _Energy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            e = module.energy(theta.detach().cpu().numpy())
            ctx.save_for_backward(theta.detach())
            return torch.tensor(e, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            (theta,) = ctx.saved_tensors
            # reverse sweep: forward state again, then the adjoint walk back
            _, g = module.energy_and_grad(theta.cpu().numpy())
            return grad_out * torch.as_tensor(g, dtype=theta.dtype, device=theta.device)

    return _Energy.apply
