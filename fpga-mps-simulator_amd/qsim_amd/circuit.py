"""A minimal circuit object with the attributes the drivers read from a Qiskit circuit.

FPGAMPSSimulator.run and FPGASimulator.run read `num_qubits`, `num_clbits`,
`name`, `qubits`, `clbits`, `find_bit(bit).index` and `data[i].operation.name /
.params`, `data[i].qubits`, `data[i].clbits`.  Qiskit is not a dependency; this
class provides exactly that surface so circuits can be built without it.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence


@dataclass(frozen=True)
class Bit:
    kind: str
    index: int


@dataclass
class Operation:
    name: str
    params: List[float] = field(default_factory=list)


@dataclass
class Instruction:
    operation: Operation
    qubits: tuple
    clbits: tuple = ()


@dataclass
class _Loc:
    index: int


class QuantumCircuit:
    def __init__(self, num_qubits: int, num_clbits: int = 0, name: str = 'circuit'):
        self.num_qubits = num_qubits
        self.num_clbits = num_clbits
        self.name = name
        self.qubits = [Bit('q', i) for i in range(num_qubits)]
        self.clbits = [Bit('c', i) for i in range(num_clbits)]
        self.data: List[Instruction] = []

    def find_bit(self, bit: Bit) -> _Loc:
        return _Loc(bit.index)

    def append(self, name: str, qubits: Sequence[int], params: Optional[Sequence[float]] = None,
               clbits: Sequence[int] = ()):
        for q in qubits:
            if not 0 <= q < self.num_qubits:
                raise ValueError(f'qubit {q} out of range')
        self.data.append(Instruction(Operation(name, [float(p) for p in (params or [])]),
                                     tuple(self.qubits[q] for q in qubits),
                                     tuple(self.clbits[c] for c in clbits)))
        return self

    def __getattr__(self, name):
        # qc.h(0), qc.cx(0, 1), qc.rz(theta, 0), qc.u(th, ph, la, 0), qc.cp(theta, 0, 1) ...
        if name.startswith('_'):
            raise AttributeError(name)
        nparam = {'rx': 1, 'ry': 1, 'rz': 1, 'p': 1, 'phase': 1, 'u1': 1, 'u2': 2, 'u3': 3,
                  'u': 3, 'cp': 1, 'cphase': 1, 'crx': 1, 'cry': 1, 'crz': 1, 'rxx': 1,
                  'ryy': 1, 'rzz': 1}.get(name, 0)

        def add(*args):
            params, qs = list(args[:nparam]), [int(a) for a in args[nparam:]]
            return self.append(name, qs, params)
        return add

    def measure(self, qubit: int, clbit: int):
        return self.append('measure', [qubit], None, [clbit])

    def measure_all(self):
        if self.num_clbits < self.num_qubits:
            self.clbits += [Bit('c', i) for i in range(self.num_clbits, self.num_qubits)]
            self.num_clbits = self.num_qubits
        for q in range(self.num_qubits):
            self.measure(q, q)
        return self

    def reset(self, qubit: int):
        return self.append('reset', [qubit])

    def barrier(self, *qubits):
        return self.append('barrier', list(qubits) or list(range(self.num_qubits)))
