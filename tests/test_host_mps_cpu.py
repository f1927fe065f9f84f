"""The MPS object the drop-in returns (FPGAMPSResult.get_mps -> HostMPS) against the
reference's MPSState read-outs: bond spectra against the golden vectors written from
mps_core.py, entanglement entropy against the oracle and against the Schmidt spectrum
of the dense state, canonical forms by invariance of the state.  Host-only (no GPU)."""

import math
import os

import numpy as np
import pytest

from oracle import mps_ref
from qsim_amd.mps import HostMPS

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')


def golden_chain(name):
    case = dict(np.load(os.path.join(GOLD, f'mps_core_{name}.npz')))
    m = mps_ref.MPS(int(case['n']), int(case['chi_max']), float(case['cutoff']))
    for g, i in zip(case['gates'], case['sites']):
        m.apply_2q(int(i), int(i) + 1, g)
    host = HostMPS([t.copy() for t in m.sites], m.norm_log, int(case['chi_max']),
                   float(case['cutoff']))
    return case, m, host


@pytest.mark.parametrize('name', ['exact_10q', 'trunc_12q', 'c1_16q'])
def test_bond_spectrum_matches_mps_core_golden(name):
    case, _, host = golden_chain(name)
    for b in range(host.n - 1):
        s = host.bond_spectrum(b)
        assert np.allclose(s, case['bond_spectra'][b][:s.size], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize('name', ['exact_10q', 'trunc_12q'])
def test_entanglement_entropy_matches_oracle_and_schmidt_spectrum(name):
    case, m, host = golden_chain(name)
    sv = case['statevector']
    n = host.n
    # the reference's formula gives the Schmidt entropy once the sites right of the bond
    # are right-normalised (its docstring asks for mixed-canonical form)
    rc = host.copy()
    rc.right_canonicalize()
    for b in range(n - 1):
        e = host.entanglement_entropy(b)
        assert abs(e - m.entanglement_entropy(b)) < 1e-10
        # Schmidt spectrum of the dense state across the cut after site b: sites 0..b are
        # qubits 0..b, the low index bits
        mat = sv.reshape(2 ** (n - b - 1), 2 ** (b + 1)).T
        s = np.linalg.svd(mat, compute_uv=False)
        p = s[s > 1e-12 * s[0]] ** 2
        p = p / p.sum()
        assert abs(rc.entanglement_entropy(b) + np.sum(p * np.log(p))) < 1e-9
    with pytest.raises(ValueError):
        host.entanglement_entropy(n - 1)


def test_canonical_forms_keep_the_state():
    _, m, host = golden_chain('exact_10q')
    sv = host.get_statevector()
    for how in ('left_canonicalize', 'right_canonicalize'):
        h = host.copy()
        getattr(h, how)()
        assert np.max(np.abs(h.get_statevector() - sv)) < 1e-12
    h = host.copy()
    h.left_canonicalize()
    for t in h.sites[:-1]:
        a = t.reshape(-1, t.shape[2])
        assert np.allclose(a.conj().T @ a, np.eye(a.shape[1]), atol=1e-12)
    h = host.copy()
    h.right_canonicalize()
    for t in h.sites[1:]:
        a = t.reshape(t.shape[0], -1)
        assert np.allclose(a @ a.conj().T, np.eye(a.shape[0]), atol=1e-12)
    # copy is independent
    h = host.copy()
    h.sites[0][...] = 0
    assert np.max(np.abs(host.get_statevector() - sv)) == 0


def test_apply_1q_matches_oracle():
    _, m, host = golden_chain('exact_10q')
    g = mps_ref.gate_1q('u3', [0.3, 1.1, -0.4])
    for q in (0, 4, 9):
        host.apply_1q(q, g)
        m.apply_1q(q, g)
    assert np.max(np.abs(host.get_statevector() - m.get_statevector())) < 1e-13
    assert host.chi_l(3) == m.sites[3].shape[0] and host.chi_r(3) == m.sites[3].shape[2]
