"""The 1000-qubit chi-64 workload at full size (brick-wall depth 40 of exp(-0.1iH),
19 980 two-site updates, cutoff 1e-10) against the oracle's own run of it
(tests/golden/c3_1000q_chi64.npz, written on the CPU by tests/golden/make_c3_fixture.py).

FP64: identical bond dimensions after every layer, norm_log at 1e-10 relative, <Z_i>
and <Z_i Z_i+1> (device transfer matrices) and the overlaps with four random product
states close to the oracle's.  FP32-MFMA contraction: the same bond dimensions, and an
infidelity of at most 1e-6 against the FP64 GPU chain that the checks above tie to the
oracle (the oracle's chain itself, 128 MiB, is not stored)."""

import math
import os

import numpy as np
import pytest

from oracle import mps_ref
from qsim_amd import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def c3():
    case = np.load(os.path.join(HERE, 'golden', 'c3_1000q_chi64.npz'))
    gates = workloads.brickwall_gates(int(case['n']), int(case['depth']), int(case['seed']))
    layers = {}
    for lay, i, g in gates:
        layers.setdefault(lay, []).append((i, g))
    return case, [layers[k] for k in sorted(layers)]


def _run(case, layers, precision):
    from qsim_amd.mps import MPSChain
    ch = MPSChain(int(case['n']), int(case['chi_max']), float(case['cutoff']),
                  precision=precision)
    bonds = []
    for batch in layers:
        ch.apply_2q_batch(batch)
        bonds.append(ch.bond_dims())
    ch.check_status()
    return ch, np.array(bonds)


@pytest.fixture(scope='module')
def fp64_chain(c3):
    case, layers = c3
    return _run(case, layers, 'fp64')


def test_fp64_full_size_matches_oracle_run(c3, fp64_chain):
    case, _ = c3
    ch, bonds = fp64_chain
    assert np.array_equal(bonds, case['bond_dims_per_layer'])
    assert abs(ch.norm_log - float(case['norm_log_per_layer'][-1])) < \
        1e-10 * abs(float(case['norm_log_per_layer'][-1]))
    dz = np.max(np.abs(ch.z_expectations() - case['z']))
    dzz = np.max(np.abs(ch.zz_expectations() - case['zz']))
    print(f'max |d<Z>| {dz:.2e}, max |d<ZZ>| {dzz:.2e}')
    assert dz < 1e-9 and dzz < 1e-9
    host = ch.to_host()
    m = mps_ref.MPS(host.n, 1, 0)
    m.sites, m.norm_log = host.sites, host.norm_log
    for k, seed in enumerate(case['product_seeds']):
        lg, ph = mps_ref.product_overlap_log(m, mps_ref.random_product_state(host.n, int(seed)))
        dl = lg - float(case['overlap_logabs'][k])
        dp = np.angle(np.exp(1j * (ph - float(case['overlap_phase'][k]))))
        print(f'product state {seed}: log|<phi|psi>| {lg:.6f}, d log {dl:.2e}, d arg {dp:.2e}')
        assert abs(dl) < 1e-8 and abs(dp) < 1e-8


def test_fp32_full_size_within_1e6_of_fp64(c3, fp64_chain):
    case, layers = c3
    ch64, _ = fp64_chain
    ch32, bonds = _run(case, layers, 'fp32')
    assert np.array_equal(bonds, case['bond_dims_per_layer'])
    infid = 1.0 - ch32.fidelity(ch64)
    dz = np.max(np.abs(ch32.z_expectations() - case['z']))
    print(f'FP32 infidelity vs FP64 {infid:.2e}, max |d<Z>| vs oracle {dz:.2e}')
    assert infid < 1e-6
    assert dz < 1e-4
