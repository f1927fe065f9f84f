"""Pins the MPS oracle (oracle/mps_ref.py) before it is trusted as the checker.

Checks: golden vectors written from the reference's mps_core.py, closed-form gate
identities, the truncation rule and the norm_log closed form, brute-force
contraction at tiny chi, routed known-answer circuits against the SV oracle, and
exact sampling marginals.
"""

import math
import os

import numpy as np
import pytest

from oracle import mps_ref, sv_ref
from qsim_amd.circuit import QuantumCircuit

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')


def load(name):
    return dict(np.load(os.path.join(GOLD, f'mps_core_{name}.npz')))


def replay(case):
    m = mps_ref.MPS(int(case['n']), int(case['chi_max']), float(case['cutoff']))
    keeps, logs = [], []
    for g, i in zip(case['gates'], case['sites']):
        m.apply_2q(int(i), int(i) + 1, g)
        keeps.append(m.sites[int(i)].shape[2])
        logs.append(m.norm_log)
    return m, np.array(keeps), np.array(logs)


@pytest.mark.parametrize('name', ['c1_16q', 'trunc_12q', 'exact_10q'])
def test_oracle_reproduces_mps_core_golden(name):
    case = load(name)
    m, keeps, logs = replay(case)
    assert list(m.bond_dims()) == list(case['bond_dims'])
    assert np.array_equal(keeps, case['keeps'])
    assert np.allclose(logs, case['norm_log_traj'], rtol=1e-12, atol=1e-12)
    n, chi = int(case['n']), int(case['chi_max'])
    for b in range(n - 1):
        s = m.bond_spectrum(b)
        ref = case['bond_spectra'][b][:s.size]
        assert np.allclose(s, ref, rtol=1e-9, atol=1e-12)
    sv = m.get_statevector()
    p = np.abs(sv) ** 2
    assert abs(p.sum() - float(case['norm_sq'])) < 1e-10
    idx = np.arange(sv.size)
    z = np.array([np.sum(p * (1 - 2 * ((idx >> q) & 1))) / p.sum() for q in range(n)])
    assert np.allclose(z, case['z_expect'], atol=1e-10)
    assert np.allclose(m.z_expectations(), case['z_expect'], atol=1e-10)
    if 'statevector' in case:
        assert np.allclose(sv, case['statevector'], atol=1e-10)


def test_gate_identities():
    g1, g2 = mps_ref.gate_1q, mps_ref.gate_2q
    for nm in ['h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx', 'sxdg']:
        u = g1(nm)
        assert np.allclose(u.conj().T @ u, np.eye(2))
    for nm, p in [('rx', [0.3]), ('ry', [1.1]), ('rz', [-0.7]), ('p', [0.4]), ('u2', [0.2, 0.9]),
                  ('u3', [0.5, -1.2, 2.0])]:
        u = g1(nm, p)
        assert np.allclose(u.conj().T @ u, np.eye(2))
    assert np.allclose(g1('sx') @ g1('sx'), g1('x'))
    assert np.allclose(g1('u2', [0, math.pi]), g1('h'))
    assert np.allclose(g1('u3', [math.pi / 2, 0, math.pi]), g1('h'))
    assert np.allclose(g1('u3', [math.pi, 0, math.pi]), g1('x'))
    assert np.allclose(g1('rx', [math.pi]), -1j * g1('x'))
    assert np.allclose(g1('rz', [0.8]), np.exp(-0.4j) * g1('p', [0.8]))
    assert np.allclose(g1('t') @ g1('t'), g1('s'))
    for nm, p in [('cx', None), ('cz', None), ('cy', None), ('swap', None), ('iswap', None),
                  ('dcx', None), ('ch', None), ('cp', [0.3]), ('crx', [0.3]), ('cry', [0.3]),
                  ('crz', [0.3]), ('csx', None), ('ecr', None), ('rxx', [0.3]), ('ryy', [0.3]),
                  ('rzz', [0.3])]:
        u = g2(nm, p)
        assert np.allclose(u.conj().T @ u, np.eye(4)), nm
    # rzz = exp(-i theta/2 Z x Z), rxx = exp(-i theta/2 X x X)
    Z, X = g1('z'), g1('x')
    th = 0.7
    assert np.allclose(g2('rzz', [th]), np.diag(np.exp(-0.5j * th * np.diag(np.kron(Z, Z)))))
    w, v = np.linalg.eigh(np.kron(X, X))
    assert np.allclose(g2('rxx', [th]), (v * np.exp(-0.5j * th * w)) @ v.conj().T)
    with pytest.raises(ValueError):
        g2('cnot')     # the reference's _gate_2q has no 'cnot' matrix either


def _state_from_sv_oracle(n, gates):
    return sv_ref.run_circuit(n, gates)


def _circ(n, ops):
    qc = QuantumCircuit(n)
    for nm, qs, ps in ops:
        qc.append(nm, qs, ps)
    return qc


DIAG_KATS = [  # fpga_diag.py:28-57
    (2, [('h', [0], None), ('cx', [0, 1], None)]),
    (4, [('h', [0], None), ('cx', [0, 1], None), ('cx', [1, 2], None), ('cx', [2, 3], None)]),
    (4, [('h', [0], None), ('cx', [0, 3], None)]),
    (1, [('u', [0], [0.7, 1.3, -0.4])]),
    (2, [('u', [0], [0.7, 1.3, -0.4]), ('u', [1], [-0.9, 0.2, 2.1]), ('cx', [0, 1], None),
         ('u', [0], [0.3, -0.6, 1.1])]),
    (4, [('u', [0], [0.7, 1.3, -0.4]), ('u', [2], [-0.9, 0.2, 2.1]), ('cx', [0, 2], None)]),
]


@pytest.mark.parametrize('n,ops', DIAG_KATS)
def test_routed_kats_match_sv_oracle(n, ops):
    gl, _ = mps_ref.compile_ops(mps_ref.circuit_ops(_circ(n, ops)), n)
    m = mps_ref.run_gate_list(n, gl, 64, 1e-12)
    sv = m.get_statevector()
    ref = _state_from_sv_oracle(n, ops)
    fid = abs(np.vdot(sv, ref)) ** 2 / (np.vdot(sv, sv).real * np.vdot(ref, ref).real)
    assert fid > 1 - 1e-13


def test_nonadjacent_apply_2q_with_reversed_order():
    rng = np.random.default_rng(1)
    n = 6
    m = mps_ref.MPS(n, 64, 1e-14)
    ops = []
    for _ in range(12):
        a, b = [int(v) for v in rng.choice(n, 2, replace=False)]
        th = float(rng.uniform(-2, 2))
        m.apply_2q(a, b, mps_ref.gate_2q('crx', [th]))
        ops.append(('crx', [a, b], [th]))
    ref = sv_ref.run_circuit(n, ops)
    sv = m.get_statevector()
    assert np.allclose(sv / np.linalg.norm(sv), ref, atol=1e-12)


def test_svd_truncate_rule_and_norm_log_closed_form():
    rng = np.random.default_rng(4)
    m = mps_ref.MPS(2, 3, 1e-3)
    s = np.array([0.8, 0.5, 0.3, 0.1, 1e-5])
    s = s / np.linalg.norm(s)                       # normalised input
    u, _ = np.linalg.qr(rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5)))
    v, _ = np.linalg.qr(rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5)))
    mat = (u * s) @ v.conj().T
    _, sq, _, sfull, dlog = m.svd_truncate(mat)
    assert sq.size == 3                             # chi_max caps 4 above-cutoff values
    eps = float(np.sum(s[3:] ** 2))
    assert abs(dlog - math.log(1 - eps)) < 1e-13    # truncation weight counted twice
    assert abs(np.sum(sq ** 4) - 1) < 1e-13         # renormalised S
    # keep rule is strict: values equal to cutoff*S[0] are dropped
    m2 = mps_ref.MPS(2, 10, 0.5)
    s2 = np.array([1.0, 0.5, 0.25])
    _, sq2, _, _, _ = m2.svd_truncate(np.diag(s2).astype(complex))
    assert sq2.size == 1


def test_theta_brute_force():
    rng = np.random.default_rng(8)
    m = mps_ref.MPS(2, 8, 1e-12)
    a = rng.standard_normal((3, 2, 4)) + 1j * rng.standard_normal((3, 2, 4))
    b = rng.standard_normal((4, 2, 5)) + 1j * rng.standard_normal((4, 2, 5))
    g = mps_ref.haar_4x4(rng)
    m.sites = [a, b]
    th = m.theta(0, g)
    brute = np.zeros((6, 10), dtype=complex)
    for aa in range(3):
        for s0 in range(2):
            for s1 in range(2):
                for d in range(5):
                    v = 0
                    for bb in range(4):
                        for p in range(2):
                            for q in range(2):
                                v += g[2 * s0 + s1, 2 * p + q] * a[aa, p, bb] * b[bb, q, d]
                    brute[2 * aa + s0, 5 * s1 + d] = v
    assert np.allclose(th, brute, atol=1e-13)


def test_sampling_marginals():
    gl, meas = mps_ref.compile_ops(mps_ref.circuit_ops(
        _circ(3, [('h', [0], None), ('cx', [0, 1], None), ('cx', [1, 2], None)])), 3)
    m = mps_ref.run_gate_list(3, gl, 16, 1e-12)
    c = m.get_counts(2000, meas, np.random.default_rng(0))
    assert set(c) <= {'000', '111'} and abs(c['000'] / 2000 - 0.5) < 0.05
    # chi^2 on a random 5-qubit state
    n = 5
    rng = np.random.default_rng(2)
    m = mps_ref.MPS(n, 32, 1e-14)
    for _ in range(3):
        for i in range(n - 1):
            m.apply_2q(i, i + 1, mps_ref.haar_4x4(rng))
    p = np.abs(m.get_statevector()) ** 2
    p /= p.sum()
    shots = 20000
    c = m.get_counts(shots, None, np.random.default_rng(3))
    obs = np.zeros(1 << n)
    for k, v in c.items():
        obs[int(k, 2)] = v
    exp = p * shots
    chi2 = np.sum((obs - exp) ** 2 / np.maximum(exp, 1e-9))
    assert chi2 < 2 * (1 << n) + 30


def test_partial_environments_and_given_envs():
    """right_environments(stop=k) holds R_i for i >= k exactly as the full recursion does
    (None below k), and sample(envs=...) with the chain's own environments draws the same
    shots as sample() from the same generator (the bench's CPU sampling leg times both)."""
    n = 8
    rng = np.random.default_rng(4)
    m = mps_ref.MPS(n, 16, 1e-14)
    for _ in range(3):
        for i in range(n - 1):
            m.apply_2q(i, i + 1, mps_ref.haar_4x4(rng))
    full = m.right_environments()
    part = m.right_environments(stop=5)
    assert all(e is None for e in part[:5])
    for i in range(5, n + 1):
        assert np.array_equal(part[i], full[i])
    assert m.sample(50, None, np.random.default_rng(9), envs=full) == \
        m.sample(50, None, np.random.default_rng(9))
    # identity environments are a valid (different) conditional model of the same shapes
    eye = [np.eye(a.shape[0], dtype=np.complex128) for a in m.sites] + [full[n]]
    shots = m.sample(20, None, np.random.default_rng(9), envs=eye)
    assert len(shots) == 20 and all(len(x) == n and set(x) <= {'0', '1'} for x in shots)


def test_pauli_expectation_closed_form():
    sv = np.zeros(8, dtype=complex)
    sv[0b011] = 1                     # qubits 0,1 = 1
    assert mps_ref.pauli_expectation_sv(sv, 'ZZZ') == 1.0    # (+1)(-1)(-1)
    assert mps_ref.pauli_expectation_sv(sv, 'ZII') == 1.0    # qubit 2 is 0
    assert mps_ref.pauli_expectation_sv(sv, 'IIZ') == -1.0
    plus = np.full(4, 0.5, dtype=complex)
    assert abs(mps_ref.pauli_expectation_sv(plus, 'XX') - 1) < 1e-15


def test_tfim_energy_matches_dense():
    n = 6
    rng = np.random.default_rng(5)
    m = mps_ref.MPS(n, 32, 1e-14)
    for _ in range(2):
        for i in range(n - 1):
            m.apply_2q(i, i + 1, mps_ref.haar_4x4(rng))
    sv = m.get_statevector()
    sv = sv / np.linalg.norm(sv)
    e = 0.0
    for i in range(n):
        e -= mps_ref.pauli_expectation_sv(sv, ''.join('X' if n - 1 - k == i else 'I' for k in range(n)))
    for i in range(n - 1):
        e -= mps_ref.pauli_expectation_sv(sv, ''.join('Z' if n - 1 - k in (i, i + 1) else 'I'
                                                      for k in range(n)))
    assert abs(mps_ref.tfim_energy(m) - e) < 1e-12


def test_product_z_expectations_match_oracle():
    """qsim_amd.mps.z_expectations (host read-out of a device chain) on random chains."""
    from qsim_amd.mps import HostMPS, z_expectations
    rng = np.random.default_rng(3)
    for dims in ([1, 2, 4, 7, 8, 8, 6, 4, 2, 1], [1, 2, 2, 1], [1, 1]):
        n = len(dims) - 1
        sites = [rng.standard_normal((dims[i], 2, dims[i + 1])) +
                 1j * rng.standard_normal((dims[i], 2, dims[i + 1])) for i in range(n)]
        m = mps_ref.MPS(n, 16, 0)
        m.sites = [s.copy() for s in sites]
        got = z_expectations(HostMPS(sites, 0.0))
        assert np.max(np.abs(got - m.z_expectations())) < 1e-12


def test_oracle_zz_and_product_overlaps_match_dense_statevector():
    """zz_neighbours, local_expectations and product_overlap_log (the quantities the
    1000-qubit fixture records) against the dense statevector at 10 qubits."""
    from qsim_amd import workloads
    n = 10
    m = mps_ref.MPS(n, 16, 1e-12)
    for _, i, g in workloads.brickwall_gates(n, 6, 3, 'haar'):
        m.apply_adjacent_2q(i, g)
    sv = m.get_statevector()
    idx = np.arange(sv.size)
    p = np.abs(sv) ** 2
    zq = lambda q: 1 - 2 * ((idx >> q) & 1)
    zz = [np.sum(p * zq(q) * zq(q + 1)) / p.sum() for q in range(n - 1)]
    z = [np.sum(p * zq(q)) / p.sum() for q in range(n)]
    assert np.max(np.abs(mps_ref.zz_neighbours(m) - zz)) < 1e-13
    assert np.max(np.abs(mps_ref.local_expectations(m, mps_ref.FIXED_GATES['z']) - z)) < 1e-13
    for seed in (1, 2):
        phi = mps_ref.random_product_state(n, seed)
        full = phi[0]
        for q in range(1, n):
            full = np.kron(phi[q], full)              # qubit q on index bit q
        ov = np.vdot(full, sv)
        lg, ph = mps_ref.product_overlap_log(m, phi)
        assert abs(lg - np.log(abs(ov))) < 1e-12 and abs(np.angle(np.exp(1j * (ph - np.angle(ov))))) < 1e-12


def test_c3_fixture_is_consistent():
    import os
    case = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden',
                                'c3_1000q_chi64.npz'))
    b = case['bond_dims_per_layer']
    assert b.shape == (40, 999) and b.max() == 64 and b.min() >= 1 and b[-1].min() >= 2
    assert np.all(np.abs(case['z']) <= 1) and np.all(np.abs(case['zz']) <= 1)
    assert np.all(np.diff(case['norm_log_per_layer']) > 0)     # truncation only adds weight
