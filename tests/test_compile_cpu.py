"""The MPS circuit compiler and shot keys against outputs the reference wrote.

tests/golden/mps_compile_ccx_6q.npz holds the reference compile_circuit's gate list
for golden/fixture_circuits.ccx_circuit(), the state its gate loop reaches and its
seeded shots; tests/golden/mps_core_samples_10q.npz holds seeded mps_core.sample
shots for the full register and measured subsets.  Host-only (no GPU)."""

import os
from collections import Counter

import numpy as np
import pytest

from golden.fixture_circuits import ccx_circuit
from oracle import mps_ref, sv_ref

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')


def _fixture():
    return np.load(os.path.join(GOLD, 'mps_compile_ccx_6q.npz'))


def _same_gate(case, k, typ, name, sites, mat):
    kind = {'1q': 0, '2q': 1, 'reset': 2}[typ]
    assert int(case['kinds'][k]) == kind, k
    assert str(case['names'][k]) == name, k
    assert [int(s) for s in case['sites'][k] if s >= 0] == list(sites), k
    if mat is not None:
        d = mat.shape[0]
        assert np.array_equal(case['mats'][k][:d, :d], mat), k


def test_dropin_compile_equals_reference_compiler():
    import fpga_mps_simulator as fms
    case = _fixture()
    gl, meas = fms.compile_circuit(ccx_circuit(), 6)
    assert len(gl) == len(case['kinds'])
    for k, g in enumerate(gl):
        _same_gate(case, k, g['type'], g['gate_name'], g['sites'], g['gate_mat'])
    assert meas == [int(q) for q in case['measured']]


def test_oracle_compile_and_gate_loop_equal_reference():
    case = _fixture()
    gl, meas = mps_ref.compile_ops(mps_ref.circuit_ops(ccx_circuit()), 6)
    assert len(gl) == len(case['kinds'])
    for k, g in enumerate(gl):
        _same_gate(case, k, g['type'], g['name'], g['sites'], g['mat'])
    m = mps_ref.run_gate_list(6, gl, int(case['chi_max']), float(case['cutoff']))
    assert np.max(np.abs(m.get_statevector() - case['statevector'])) < 1e-12
    assert m.bond_dims() == [int(b) for b in case['bond_dims']]
    got = m.sample(int(case['shots']), meas, np.random.default_rng(int(case['rng_seed'])))
    assert got == [str(s) for s in case['samples']]


def test_fixed_ccx_orientation_is_opt_in_and_exact():
    """With fix_ccx_orientation the adjacent Toffoli sub-CX gates get the reversed-qubit
    matrix, and the compiled circuit equals the exact Toffoli on the statevector oracle;
    the reference default differs from the exact circuit (its control/target swap)."""
    import fpga_mps_simulator as fms
    from qsim_amd.circuit import QuantumCircuit
    rng = np.random.default_rng(5)
    qc = QuantumCircuit(4)
    ops = []
    for q in range(4):
        ps = [float(v) for v in rng.uniform(0, 6.28, 3)]
        qc.u(*ps, q)
        ops.append(('u3', [q], ps))
    qc.ccx(3, 2, 1)
    ops.append(('ccx', [3, 2, 1], None))
    exact = sv_ref.run_circuit(4, ops)

    def run(gl):
        m = mps_ref.run_gate_list(4, [{'type': g['type'], 'sites': g['sites'],
                                       'mat': g['gate_mat']} for g in gl], 64, 1e-12)
        return m.get_statevector()

    ref_gl, _ = fms.compile_circuit(qc, 4)
    fix_gl, _ = fms.compile_circuit(qc, 4, fix_ccx_orientation=True)
    differ = [k for k, (a, b) in enumerate(zip(ref_gl, fix_gl))
              if a['gate_mat'] is not None and not np.array_equal(a['gate_mat'], b['gate_mat'])]
    assert differ and all(ref_gl[k]['gate_name'] == 'cx' for k in differ)
    fid = lambda a, b: abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
    assert fid(run(fix_gl), exact) > 1 - 1e-12
    assert fid(run(ref_gl), exact) < 1 - 1e-3


def test_cnot_raises_and_unknown_names_are_skipped():
    import fpga_mps_simulator as fms
    from qsim_amd.circuit import QuantumCircuit
    qc = QuantumCircuit(3)
    qc.append('cnot', [0, 1])
    with pytest.raises(ValueError):
        fms.compile_circuit(qc, 3)
    qc = QuantumCircuit(3)
    qc.h(0)
    qc.append('no_such_gate', [1])
    qc.cx(0, 1)
    gl, meas = fms.compile_circuit(qc, 3)
    assert [g['gate_name'] for g in gl] == ['h', 'cx'] and meas == [0, 1, 2]


@pytest.mark.parametrize('name', ['full', 'sub_a', 'sub_b', 'sub_c'])
def test_oracle_samples_equal_mps_core_fixture(name):
    case = np.load(os.path.join(GOLD, 'mps_core_samples_10q.npz'))
    m = mps_ref.MPS(int(case['n']), int(case['chi_max']), float(case['cutoff']))
    for g, i in zip(case['gates'], case['sites']):
        m.apply_adjacent_2q(int(i), g)
    meas = [int(q) for q in case[f'{name}_measured']] or None
    exp = [str(s) for s in case[f'{name}_samples']]
    got = m.sample(len(exp), meas, np.random.default_rng(int(case[f'{name}_rng_seed'])))
    assert got == exp
    if meas is not None and sorted(meas) != sorted(int(case['n']) - 1 - q for q in meas):
        # the key takes the mirrored site's outcome: with the site's own outcome the
        # counts of a non-mirror-symmetric subset come out different
        own = []
        for s_full in m.sample(len(exp), None, np.random.default_rng(int(case[f'{name}_rng_seed']))):
            bits = s_full[::-1]
            own.append(''.join(bits[q] for q in sorted(meas, reverse=True)))
        assert Counter(own) != Counter(exp)


def test_dropin_pauli_expectation_matches_dense_kronecker():
    """The drop-in's _pauli_expectation_sv (fpga_mps_simulator.py:1091-1108: character k
    acts on qubit n-1-k) against the explicit Kronecker product and the oracle."""
    import fpga_mps_simulator as fms
    rng = np.random.default_rng(3)
    n = 5
    sv = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    sv /= np.linalg.norm(sv)
    mats = {'I': np.eye(2), 'X': np.array([[0, 1], [1, 0]]),
            'Y': np.array([[0, -1j], [1j, 0]]), 'Z': np.diag([1, -1])}
    for _ in range(20):
        p = ''.join(rng.choice(list('IXYZ'), n))
        full = np.array([[1.0]])
        for ch in p:                      # first character = most significant = qubit n-1
            full = np.kron(full, mats[ch])
        want = float(np.real(np.vdot(sv, full @ sv)))
        assert abs(fms._pauli_expectation_sv(sv, p) - want) < 1e-12
        assert abs(mps_ref.pauli_expectation_sv(sv, p) - want) < 1e-12
