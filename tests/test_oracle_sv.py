"""Pins the statevector oracle (oracle/sv_ref.py) before it is trusted as the checker.

Kernel matrices against closed forms and the MPS gate table, pair updates against
dense Kronecker products, the lowering against exact gate matrices up to the
documented global phase, and run_circuit against comprehensive_test_v07.py KATs.
"""

import math

import numpy as np
import pytest

from oracle import mps_ref, sv_ref

OPS = sv_ref.KERNEL_GATE_OPS


def dense_1q(n, t, m):
    out = np.array([[1.0 + 0j]])
    for q in range(n - 1, -1, -1):
        out = np.kron(out, m if q == t else np.eye(2))
    return out


def dense_perm(n, fn):
    u = np.zeros((1 << n, 1 << n), dtype=complex)
    for i in range(1 << n):
        j, ph = fn(i)
        u[j, i] = ph
    return u


@pytest.mark.parametrize('name', ['h', 'x', 'y', 'z', 's', 't', 'sdg', 'tdg', 'sx', 'sxdg'])
def test_fixed_kernel_matrices_equal_gate_table(name):
    assert np.allclose(sv_ref.kernel_matrix(OPS[name]), mps_ref.gate_1q(name), atol=1e-15)


@pytest.mark.parametrize('name,p', [('rx', [0.3]), ('ry', [-1.2]), ('rz', [2.5]), ('p', [0.9]),
                                    ('u2', [0.4, -0.8]), ('u3', [1.3, 0.2, -2.2])])
def test_param_kernel_matrices_equal_gate_table(name, p):
    pv = p + [0.0] * (3 - len(p))
    assert np.allclose(sv_ref.kernel_matrix(OPS[name], *pv), mps_ref.gate_1q(name, p), atol=1e-15)


def test_pair_updates_match_dense():
    rng = np.random.default_rng(0)
    n = 5
    for t in range(n):
        st = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
        m = mps_ref.gate_1q('u3', [0.3, 1.0, -0.5])
        ref = dense_1q(n, t, m) @ st
        sv_ref.apply_1q(st, t, m)
        assert np.allclose(st, ref)
    for c in range(n):
        for t in range(n):
            if c == t:
                continue
            st = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
            ucx = dense_perm(n, lambda i: (i ^ (1 << t) if (i >> c) & 1 else i, 1.0))
            ucz = dense_perm(n, lambda i: (i, -1.0 if (i >> c) & (i >> t) & 1 else 1.0))
            a, b = st.copy(), st.copy()
            sv_ref.apply_cx(a, n, c, t)
            sv_ref.apply_cz(b, n, c, t)
            assert np.allclose(a, ucx @ st)
            assert np.allclose(b, ucz @ st)


def lowered_unitary(n, gates):
    u = np.zeros((1 << n, 1 << n), dtype=complex)
    for k in range(1 << n):
        st = np.zeros(1 << n, dtype=complex)
        st[k] = 1
        sv_ref.apply_kernel_ops(st, n, sv_ref.lower(gates))
        u[:, k] = st
    return u


def exact_2q(g, a, b, n=2):
    u = np.zeros((1 << n, 1 << n), dtype=complex)
    for i in range(1 << n):
        col = 2 * ((i >> a) & 1) + ((i >> b) & 1)
        for row in range(4):
            j = (i & ~(1 << a) & ~(1 << b)) | (((row >> 1) & 1) << a) | ((row & 1) << b)
            u[j, i] += g[row, col]
    return u


@pytest.mark.parametrize('name,p', [('cy', None), ('ch', None), ('swap', None), ('crx', [0.7]),
                                    ('cry', [0.7]), ('crz', [0.7]), ('csx', None),
                                    ('iswap', None), ('ecr', None), ('rxx', [0.7]),
                                    ('ryy', [0.7]), ('rzz', [0.7])])
def test_lowering_exact(name, p):
    for a, b in ((0, 1), (1, 0)):
        lu = lowered_unitary(2, [(name, [a, b], p)])
        assert np.allclose(lu, exact_2q(mps_ref.gate_2q(name, p), a, b), atol=1e-14)


def test_lowering_cp_phase_and_dcx_order():
    th = 0.7
    lu = lowered_unitary(2, [('cp', [0, 1], [th])])
    assert np.allclose(lu, np.exp(0.25j * th) * exact_2q(mps_ref.gate_2q('cp', [th]), 0, 1))
    # SV dcx(a,b) = cx(a,b) then cx(b,a) = the transpose of the MPS table's dcx
    lu = lowered_unitary(2, [('dcx', [0, 1], None)])
    assert np.allclose(lu, exact_2q(mps_ref.gate_2q('dcx').T, 0, 1))
    # Toffoli decomposition is exact
    lu = lowered_unitary(3, [('ccx', [0, 1, 2], None)])
    assert np.allclose(lu, dense_perm(3, lambda i: (i ^ 4 if (i & 3) == 3 else i, 1.0)), atol=1e-14)


@pytest.mark.parametrize('nq,gates,det', [
    (1, [('x', [0], None)], 1),
    (2, [('x', [0], None), ('cx', [0, 1], None)], 3),
    (2, [('x', [0], None), ('swap', [0, 1], None)], 2),
    (3, [('x', [0], None), ('x', [1], None), ('ccx', [0, 1, 2], None)], 7),
    (1, [('sx', [0], None), ('sx', [0], None)], 1),
    (2, [('x', [0], None), ('iswap', [0, 1], None)], 2),
])
def test_kats(nq, gates, det):
    st = sv_ref.run_circuit(nq, gates)
    assert abs(abs(st[det]) - 1) < 1e-14


def test_qft_uniform_and_roundtrip():
    n = 6
    st = sv_ref.run_circuit(n, sv_ref.qft_gates(n))
    assert np.allclose(np.abs(st), 2 ** (-n / 2))
    st = sv_ref.run_circuit(n, sv_ref.qft_gates(n) + sv_ref.qft_gates(n, inverse=True))
    assert abs(abs(st[0]) - 1) < 1e-13


def test_reset_fallback_and_clbit_keys():
    st = sv_ref.run_circuit(1, [('x', [0], None), ('reset', [0], None)])
    assert abs(st[0] - 1) < 1e-15
    assert sv_ref.clbit_key(1, 1, 3, [(0, 0), (0, 1)]) == '011'
    assert sv_ref.clbit_key(1, 1, 3, [(0, 2)]) == '100'


def test_openmp_pair_updates_match_numpy_restatement():
    """oracle/sv_kernel_omp.c (the C/OpenMP CPU baseline) against the NumPy pair updates
    on the lowered benchmark structure and random gates."""
    from oracle import sv_omp
    from qsim_amd import workloads
    nq = 12
    rng = np.random.default_rng(4)
    gates = workloads.u3_cz_layers(nq, 3, seed=2) + workloads.qft_gates(nq)
    gates += [('cx', [int(a) for a in rng.choice(nq, 2, replace=False)]) for _ in range(20)]
    ops = sv_ref.lower(gates)
    ref = np.zeros(1 << nq, dtype=np.complex128)
    ref[0] = 1.0
    sv_ref.apply_kernel_ops(ref, nq, ops)
    re = np.zeros(1 << nq)
    im = np.zeros(1 << nq)
    re[0] = 1.0
    sv_omp.apply_kernel_ops(re, im, nq, ops)
    assert np.max(np.abs(re + 1j * im - ref)) < 1e-13
    assert sv_omp.threads() >= 1
