"""Device transfer-matrix read-outs (qs_mps_left_envs, qs_mps_expect, qs_mps_overlap)
against the MPS oracle's transfer matrices and, at small n, the dense statevector."""

import math

import numpy as np
import pytest

from oracle import mps_ref
from qsim_amd import workloads

pytestmark = pytest.mark.gpu

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1.0, -1.0]).astype(complex)


def _chain(n, chi, depth, seed, kind='haar', precision='fp64'):
    from qsim_amd.mps import MPSChain
    ch = MPSChain(n, chi, 1e-12, precision=precision)
    ref = mps_ref.MPS(n, chi, 1e-12)
    batch, cur = [], None
    for lay, i, g in workloads.brickwall_gates(n, depth, seed=seed, kind=kind):
        if lay != cur and batch:
            ch.apply_2q_batch(batch)
            batch = []
        cur = lay
        batch.append((i, g))
    ch.apply_2q_batch(batch)
    ch.check_status()
    ref.sites = [t.copy() for t in ch.to_host().sites]      # the device chain, on the host
    ref.norm_log = ch.norm_log
    return ch, ref


@pytest.mark.parametrize('n,chi,depth', [(12, 16, 8), (30, 64, 14), (14, 100, 16)])
def test_left_environments_and_expectations_match_oracle(n, chi, depth):
    ch, ref = _chain(n, chi, depth, seed=n + chi)
    left, _ = ch.environments()
    want, _ = mps_ref._environments(ref)
    L = left.cpu().numpy().reshape(n + 1, chi, chi)
    b = [1] + ref.bond_dims() + [1]
    for i in range(n + 1):
        d = b[i]
        assert np.allclose(L[i, :d, :d], want[i], rtol=1e-10, atol=1e-13), i
    for op in (Z, X, Y):
        got = ch.expectations(op)
        assert np.max(np.abs(got - mps_ref.local_expectations(ref, op))) < 1e-10
    assert np.max(np.abs(ch.zz_expectations() - mps_ref.zz_neighbours(ref))) < 1e-10


def test_pair_expectations_match_dense_statevector():
    n = 10
    ch, ref = _chain(n, 32, 6, seed=4)
    sv = ref.get_statevector()
    p = sv.reshape([2] * n)                    # axis k is qubit n-1-k
    got = ch.pair_expectations(X, Z)
    for i in range(n - 1):
        t = np.tensordot(Z, p, axes=([1], [n - 2 - i]))
        t = np.moveaxis(t, 0, n - 2 - i)
        t = np.tensordot(X, t, axes=([1], [n - 1 - i]))
        t = np.moveaxis(t, 0, n - 1 - i)
        want = np.real(np.vdot(sv, t.reshape(-1))) / np.real(np.vdot(sv, sv))
        assert abs(got[i] - want) < 1e-10, i


def test_overlap_and_fidelity_match_oracle():
    n, chi = 16, 32
    a, ra = _chain(n, chi, 8, seed=1)
    b, rb = _chain(n, chi, 8, seed=1, precision='fp32')      # FP32 contraction: nearby
    c, rc = _chain(n, chi, 8, seed=2)
    for x, rx, y, ry in ((a, ra, b, rb), (a, ra, c, rc), (c, rc, c, rc)):
        lg, ph = x.overlap_log(y)
        want = mps_ref.overlap(rx, ry)
        assert abs(lg - math.log(abs(want))) < 1e-10
        assert abs(np.angle(np.exp(1j * (ph - np.angle(want))))) < 1e-9
    f = a.fidelity(b)
    fw = abs(mps_ref.overlap(ra, rb)) ** 2 / (mps_ref.overlap(ra, ra).real *
                                                mps_ref.overlap(rb, rb).real)
    assert abs(f - fw) < 1e-12 and 1 - f < 1e-6


def test_thousand_site_chain_readouts_stay_finite():
    n = 1000
    a, ra = _chain(n, 16, 6, seed=7, kind='near_identity')
    z = a.z_expectations()
    assert np.all(np.isfinite(z)) and np.all(np.abs(z) <= 1 + 1e-12)
    assert np.max(np.abs(z - mps_ref.local_expectations(ra, Z))) < 1e-10
    lg, _ = a.overlap_log(a)
    assert math.isfinite(lg)
    assert abs(a.fidelity(a) - 1.0) < 1e-12


@pytest.mark.parametrize('name', ['exact_10q', 'trunc_12q', 'c1_16q'])
def test_device_entropy_and_spectrum_match_golden_chains(name):
    """Entanglement entropy on the device (left Gram environments + Jacobi eigenvalues)
    and bond_spectrum from the device bond dimensions, on the chains of the mps_core
    golden vectors, against the oracle's restatement of mps_core (dense contraction +
    LAPACK) and the recorded spectra."""
    import os
    from qsim_amd.mps import MPSChain
    here = os.path.dirname(os.path.abspath(__file__))
    case = np.load(os.path.join(here, 'golden', f'mps_core_{name}.npz'))
    n = int(case['n'])
    ch = MPSChain(n, int(case['chi_max']), float(case['cutoff']))
    ref = mps_ref.MPS(n, int(case['chi_max']), float(case['cutoff']))
    for g, i in zip(case['gates'], case['sites']):
        ch.apply_2q(int(i), g)
        ref.apply_2q(int(i), int(i) + 1, g)
    ch.check_status()
    got = ch.entanglement_entropies()
    for b in range(n - 1):
        assert abs(got[b] - ref.entanglement_entropy(b)) < 1e-10, b
        s = ch.bond_spectrum(b)
        assert np.allclose(s, case['bond_spectra'][b][:s.size], rtol=1e-9, atol=1e-12)
        assert s.size == np.count_nonzero(case['bond_spectra'][b])


def test_device_entropy_chi100_class():
    ch, ref = _chain(14, 100, 14, seed=3)
    got = ch.entanglement_entropies([0, 3, 6, 7, 12])
    for b, e in zip([0, 3, 6, 7, 12], got):
        assert abs(e - ref.entanglement_entropy(b)) < 1e-10, b
