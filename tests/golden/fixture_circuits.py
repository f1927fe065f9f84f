"""Circuits whose reference outputs are recorded as fixtures (inputs only).

Built with the repository's duck-typed circuit class, so the same object can be
handed to the reference's compiler (when the fixtures are written) and to the
drop-in (when they are checked).
"""

import math

import numpy as np

from qsim_amd.circuit import QuantumCircuit


def ccx_circuit() -> QuantumCircuit:
    """6 qubits: Toffolis whose adjacent sub-CX gates have the control site above the
    target site, a routed two-qubit gate, a reset, an unknown gate name and a measured
    subset."""
    rng = np.random.default_rng(77)
    qc = QuantumCircuit(6, 6)
    for q in range(6):
        qc.u(*[float(v) for v in rng.uniform(0, 2 * math.pi, 3)], q)
    qc.cx(0, 1)
    qc.ccx(3, 2, 1)          # cx(2,1) and cx(3,2): control site above the target
    qc.ry(0.37, 4)
    qc.ccx(0, 1, 2)          # every sub-CX has its control below its target
    qc.ccx(5, 3, 4)          # cx(3,4) below, cx(5,4) above, cx(5,3) routed
    qc.append('unknown_op', [2])
    qc.cp(0.91, 0, 4)        # routed
    qc.reset(3)
    qc.h(3)
    qc.rzz(0.55, 2, 3)
    for q in (1, 4, 5):
        qc.measure(q, q)
    return qc
