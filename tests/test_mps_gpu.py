"""GPU parity tests of the MPS path (libqsim_hip.so) against the MPS oracle.

FP64 GPU runs must match mps_core semantics to 1e-10 with identical bond dimensions
(golden vectors written from mps_core.py, and the oracle on larger inputs); the
complex-FP32 MFMA contraction must stay within 1e-6 infidelity of the FP64 oracle.
"""

import math
import os

import numpy as np
import pytest

from oracle import mps_ref, sv_ref
from qsim_amd import workloads
from qsim_amd.circuit import QuantumCircuit

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def chain_run(n, chi, cutoff, updates, precision='fp64'):
    """updates: [(layer, site, gate)] -> MPSChain after batching each layer."""
    from qsim_amd.mps import MPSChain
    ch = MPSChain(n, chi, cutoff, precision=precision)
    cur, batch = None, []
    for lay, i, g in updates:
        if lay != cur and batch:
            ch.apply_2q_batch(batch)
            batch = []
        cur = lay
        batch.append((i, g))
    ch.apply_2q_batch(batch)
    ch.check_status()
    return ch


def as_oracle(host):
    m = mps_ref.MPS(host.n, 1, 0)
    m.sites = host.sites
    m.norm_log = host.norm_log
    return m


def fidelity(a, b):
    ab = mps_ref.overlap(a, b)
    aa = mps_ref.overlap(a, a).real
    bb = mps_ref.overlap(b, b).real
    return abs(ab) ** 2 / (aa * bb)


@pytest.mark.parametrize('name', ['c1_16q', 'trunc_12q', 'exact_10q'])
def test_golden_vectors_fp64(name):
    case = dict(np.load(os.path.join(HERE, 'golden', f'mps_core_{name}.npz')))
    n, chi, cut = int(case['n']), int(case['chi_max']), float(case['cutoff'])
    layers = workloads.brickwall_bonds(n, int(case['depth']))
    lay_of = {}
    k = 0
    ups = []
    for lay, bonds in enumerate(layers):
        for i in bonds:
            assert int(case['sites'][k]) == i
            ups.append((lay, i, case['gates'][k]))
            k += 1
    ch = chain_run(n, chi, cut, ups)
    host = ch.to_host()
    # the dense contraction on the device equals the host one
    if n <= 24:
        assert np.max(np.abs(ch.statevector() - host.get_statevector())) < 1e-12
    assert host.bond_dims() == [int(v) for v in case['bond_dims']]
    assert abs(host.norm_log - float(case['norm_log'])) < 1e-10 * max(1, abs(float(case['norm_log'])))
    sv = host.get_statevector()
    p = np.abs(sv) ** 2
    assert abs(p.sum() - float(case['norm_sq'])) < 1e-9
    idx = np.arange(sv.size)
    z = np.array([np.sum(p * (1 - 2 * ((idx >> q) & 1))) / p.sum() for q in range(n)])
    assert np.allclose(z, case['z_expect'], atol=1e-10)
    orc = as_oracle(host)
    for b in range(n - 1):
        s = orc.bond_spectrum(b)
        assert np.allclose(s, case['bond_spectra'][b][:s.size], rtol=1e-8, atol=1e-11)
    if 'statevector' in case:
        assert np.max(np.abs(sv - case['statevector'])) < 1e-10


def test_near_identity_chain_fp64_matches_oracle():
    n, depth, chi, cut = 48, 14, 64, 1e-10
    ups = workloads.brickwall_gates(n, depth, seed=17)
    ch = chain_run(n, chi, cut, ups)
    ref = mps_ref.MPS(n, chi, cut)
    for _, i, g in ups:
        ref.apply_adjacent_2q(i, g)
    host = ch.to_host()
    assert host.bond_dims() == ref.bond_dims()
    assert abs(host.norm_log - ref.norm_log) < 1e-10 * max(1.0, abs(ref.norm_log))
    assert 1 - fidelity(as_oracle(host), ref) < 1e-10


def test_haar_saturated_chi64_fp64_matches_oracle():
    n, depth, chi, cut = 16, 14, 64, 1e-12
    ups = workloads.brickwall_gates(n, depth, seed=5, kind='haar')
    ch = chain_run(n, chi, cut, ups)
    ref = mps_ref.MPS(n, chi, cut)
    for _, i, g in ups:
        ref.apply_adjacent_2q(i, g)
    host = ch.to_host()
    assert host.bond_dims() == ref.bond_dims()
    assert max(host.bond_dims()) == 64
    assert abs(host.norm_log - ref.norm_log) < 1e-10 * max(1.0, abs(ref.norm_log))
    sv, rv = host.get_statevector(), ref.get_statevector()
    assert np.max(np.abs(sv - rv)) < 1e-10 * np.max(np.abs(rv))


@pytest.mark.parametrize('chi,precision', [(128, 'fp64'), (100, 'fp64'), (128, 'fp32')])
def test_haar_chi128_class_matches_oracle(chi, precision):
    # chi_max > 64 selects the SVD class that keeps theta's columns in the workspace;
    # 16 qubits of Haar gates reach bond 256 untruncated, so chi 100 / 128 truncate.
    n, depth, cut = 16, 16, 1e-12
    ups = workloads.brickwall_gates(n, depth, seed=8, kind='haar')
    ch = chain_run(n, chi, cut, ups, precision=precision)
    ref = mps_ref.MPS(n, chi, cut)
    for _, i, g in ups:
        ref.apply_adjacent_2q(i, g)
    host = ch.to_host()
    assert host.bond_dims() == ref.bond_dims()
    assert max(host.bond_dims()) == chi
    sv, rv = host.get_statevector(), ref.get_statevector()
    if precision == 'fp64':
        assert abs(host.norm_log - ref.norm_log) < 1e-10 * max(1.0, abs(ref.norm_log))
        assert np.max(np.abs(sv - rv)) < 1e-10 * np.max(np.abs(rv))
    else:
        # complex-FP32 contraction on MFMA (the chi 65..128 class's block kernel)
        fid = abs(np.vdot(sv, rv)) ** 2 / (np.vdot(sv, sv).real * np.vdot(rv, rv).real)
        assert 1 - fid < 1e-6
        assert abs(host.norm_log - ref.norm_log) < 1e-6 * max(1.0, abs(ref.norm_log))


@pytest.mark.parametrize('chi', [8, 16, 32, 64])
def test_every_svd_team_width_matches_oracle(chi):
    # each built tile shape of the register SVD against the oracle: Haar gates on 14
    # qubits grow the middle bonds to 128, so every class truncates
    from qsim_amd import mps_tune
    n, depth, cut = 14, 14, 1e-12
    ups = workloads.brickwall_gates(n, depth, seed=31 + chi, kind='haar')
    ref = mps_ref.MPS(n, chi, cut)
    for _, i, g in ups:
        ref.apply_adjacent_2q(i, g)
    rv = ref.get_statevector()
    saved = mps_tune.current_lanes(chi)
    try:
        for lanes in mps_tune.lane_choices(chi):
            mps_tune.set_lanes(chi, lanes)
            host = chain_run(n, chi, cut, ups).to_host()
            assert host.bond_dims() == ref.bond_dims(), lanes
            assert abs(host.norm_log - ref.norm_log) < 1e-10 * max(1.0, abs(ref.norm_log))
            sv = host.get_statevector()
            assert np.max(np.abs(sv - rv)) < 1e-10 * np.max(np.abs(rv)), lanes
    finally:
        mps_tune.set_lanes(chi, saved)


def test_autotuner_times_every_width_and_installs_the_fastest(tmp_path):
    from qsim_amd import mps_tune
    saved = {c: mps_tune.current_lanes(c) for c in mps_tune.CLASSES}
    try:
        rep = mps_tune.autotune(classes=(16, 64), ntasks={16: 256, 64: 64}, reps=1)
        for c in (16, 64):
            times = rep[c]['ms_per_update']
            assert sorted(times) == mps_tune.lane_choices(c)
            assert all(t > 0 for t in times.values())
            assert rep[c]['lanes'] == min(times, key=times.get) == mps_tune.current_lanes(c)
        path = str(tmp_path / 'tune.json')
        mps_tune.save(path)
        mps_tune.set_lanes(16, 4 if rep[16]['lanes'] != 4 else 8)
        assert mps_tune.load(path)
        assert mps_tune.current_lanes(16) == rep[16]['lanes']
    finally:
        for c, w in saved.items():
            mps_tune.set_lanes(c, w)


def test_fp32_mfma_fidelity():
    n, depth, chi, cut = 64, 16, 64, 1e-10
    ups = workloads.brickwall_gates(n, depth, seed=23)
    ch = chain_run(n, chi, cut, ups, precision='fp32')
    ref = mps_ref.MPS(n, chi, cut)
    for _, i, g in ups:
        ref.apply_adjacent_2q(i, g)
    host = ch.to_host()
    infid = 1 - fidelity(as_oracle(host), ref)
    assert infid < 1e-6, infid
    assert abs(host.norm_log - ref.norm_log) < 1e-5 * max(1.0, abs(ref.norm_log))


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


@pytest.fixture(scope='module')
def sim():
    from fpga_mps_simulator import FPGAMPSSimulator
    return FPGAMPSSimulator(num_cards=1, chi_max=64, engine='fpga', mode='shots')


@pytest.mark.parametrize('n,ops', DIAG_KATS)
def test_diag_kats(sim, n, ops):
    res = sim.run(_circ(n, ops), shots=256, seed=1)
    sv = res.get_statevector()
    assert res.get_mps() is not None and res.get_mps().n == sv.size.bit_length() - 1
    assert np.max(np.abs(res.get_mps().get_statevector() - sv)) < 1e-12
    ref = sv_ref.run_circuit(n, ops)
    fid = abs(np.vdot(sv, ref)) ** 2 / (np.vdot(sv, sv).real * np.vdot(ref, ref).real)
    assert fid > 1 - 1e-12
    assert res.mode == 'shots_fpga'
    assert set(res.get_phase_times()) >= {'fpga_kernel', 'hbm_transfer', 'cpu_svd', 'cpu_qr',
                                          'cpu_sampling', 'host_other'}


def test_ghz3_self_test(sim):
    qc = QuantumCircuit(3)
    qc.h(0)
    qc.cx(0, 1)
    qc.cx(1, 2)
    qc.measure_all()
    res = sim.run(qc, shots=1024, seed=0)
    c = res.get_counts()
    assert (c.get('000', 0) + c.get('111', 0)) / 1024 > 0.95
    assert res.get_bond_dims() == [2, 2] and res.get_max_bond() == 2


def test_random_named_circuits_match_oracle(sim):
    rng = np.random.default_rng(31)
    names1 = ['h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx', 'sxdg', 'rx', 'ry', 'rz', 'p',
              'u2', 'u3']
    names2 = ['cx', 'cz', 'cy', 'ch', 'swap', 'cp', 'crx', 'cry', 'crz', 'csx', 'iswap', 'dcx',
              'ecr', 'rxx', 'ryy', 'rzz']
    npar = {'rx': 1, 'ry': 1, 'rz': 1, 'p': 1, 'u2': 2, 'u3': 3, 'cp': 1, 'crx': 1, 'cry': 1,
            'crz': 1, 'rxx': 1, 'ryy': 1, 'rzz': 1}
    for trial in range(4):
        n = 7
        ops = []
        for _ in range(60):
            if rng.random() < 0.5:
                nm = names1[rng.integers(len(names1))]
                qs = [int(rng.integers(n))]
            else:
                nm = names2[rng.integers(len(names2))]
                qs = [int(v) for v in rng.choice(n, 2, replace=False)]
            ps = [float(v) for v in rng.uniform(-3, 3, npar.get(nm, 0))] or None
            ops.append((nm, qs, ps))
        if trial == 3:
            ops.insert(20, ('reset', [2], None))
        qc = _circ(n, ops)
        res = sim.run(qc, shots=0, seed=3)
        gl, _ = mps_ref.compile_ops(mps_ref.circuit_ops(qc), n)
        ref = mps_ref.run_gate_list(n, gl, 64, 1e-12)
        assert res.get_bond_dims() == ref.bond_dims()
        rv = ref.get_statevector()
        assert np.max(np.abs(res.get_statevector() - rv)) < 1e-10 * max(1, np.max(np.abs(rv)))


def test_chi_limit_raises():
    from fpga_mps_simulator import FPGAMPSSimulator
    with pytest.raises(ValueError):
        FPGAMPSSimulator(chi_max=129)
    with pytest.raises(ValueError):
        FPGAMPSSimulator(engine='cpu')


# ---------------------------------------------------------------- shot sampling
def _oracle_chain(n, chi, cut, ups):
    ref = mps_ref.MPS(n, chi, cut)
    for _, i, g in ups:
        ref.apply_adjacent_2q(i, g)
    return ref


@pytest.mark.parametrize('n,chi,depth,seed', [(10, 16, 6, 3), (14, 8, 10, 4)])
def test_gpu_sampling_reproduces_oracle_samples(n, chi, depth, seed):
    """Same generator seed -> the same bitstrings as sequential sampling on the host
    (mps_core.py:371-420), for the full register and for a measured subset."""
    ups = workloads.brickwall_gates(n, depth, seed=seed, kind='haar')
    ch = chain_run(n, chi, 1e-12, ups)
    ref = _oracle_chain(n, chi, 1e-12, ups)
    got = ch.sample(3000, None, np.random.default_rng(11))
    exp = ref.sample(3000, None, np.random.default_rng(11))
    assert got == exp
    meas = [n - 1, 2, 0, 5]
    got = ch.sample(500, meas, np.random.default_rng(12))
    exp = ref.sample(500, meas, np.random.default_rng(12))
    assert got == exp and all(len(b) == 4 for b in got)


def test_host_copy_samples_and_readouts_match_oracle():
    """The drop-in's get_mps() object: shots drawn through a device copy, and the
    bond read-outs, against the oracle on the same chain."""
    n, chi = 10, 16
    ups = workloads.brickwall_gates(n, 6, seed=4, kind='haar')
    host = chain_run(n, chi, 1e-12, ups).to_host()
    ref = mps_ref.MPS(n, chi, 1e-12)
    ref.sites = [t.copy() for t in host.sites]
    ref.norm_log = host.norm_log
    assert host.sample(800, None, np.random.default_rng(3)) == \
        ref.sample(800, None, np.random.default_rng(3))
    assert host.get_counts(300, [1, 7], np.random.default_rng(4)) == \
        ref.get_counts(300, [1, 7], np.random.default_rng(4))
    for b in range(n - 1):
        assert abs(host.entanglement_entropy(b) - ref.entanglement_entropy(b)) < 1e-10
        assert np.allclose(host.bond_spectrum(b), ref.bond_spectrum(b), rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize('n,chi,depth', [(1, 2, 0), (2, 2, 1), (13, 16, 8), (20, 32, 10)])
def test_device_statevector_matches_host_contraction_and_oracle(n, chi, depth):
    from qsim_amd.mps import MPSChain
    if depth == 0:
        ch = MPSChain(n, chi, 1e-12)
        ch.apply_1q(0, np.array([[0.6, 0.8], [0.8, -0.6]], dtype=complex))
        ref = np.array([0.6, 0.8], dtype=complex)
    else:
        ups = workloads.brickwall_gates(n, depth, seed=9, kind='haar')
        ch = chain_run(n, chi, 1e-12, ups)
        ref = _oracle_chain(n, chi, 1e-12, ups).get_statevector()
    sv = ch.statevector()
    assert sv.shape == (1 << n,)
    assert np.max(np.abs(sv - ch.to_host().get_statevector())) < 1e-12
    assert np.max(np.abs(sv - ref)) < 1e-10


def test_gpu_sampling_ghz():
    from qsim_amd.mps import MPSChain
    n = 12
    ch = MPSChain(n, 4, 1e-12)
    h = np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2)
    ch.apply_1q(0, h)
    cx = np.eye(4, dtype=complex)[[0, 1, 3, 2]]        # control = left site
    for i in range(n - 1):
        ch.apply_2q(i, cx)
    c = ch.get_counts(4000, None, np.random.default_rng(5))
    assert set(c) == {'0' * n, '1' * n}
    assert abs(c['0' * n] - 2000) < 200


def test_gpu_sampling_200_qubits_chi32_matches_oracle():
    n, chi, depth = 200, 32, 12
    ups = workloads.brickwall_gates(n, depth, seed=21)
    ch = chain_run(n, chi, 1e-10, ups)
    ref = mps_ref.MPS(n, chi, 1e-10)
    ref.sites = [t.copy() for t in ch.to_host().sites]     # the device chain, on the host
    got = ch.sample(48, None, np.random.default_rng(2))
    exp = ref.sample(48, None, np.random.default_rng(2))
    assert got == exp


# ------------------------------------------- fixtures written by the reference itself
def test_gpu_samples_equal_mps_core_fixture():
    """Seeded shots of mps_core.MPSState.sample (tests/golden/mps_core_samples_10q.npz,
    written by the reference) for the full register and three measured subsets; the
    subset key takes the mirrored site's outcome (mps_core.py:415-417)."""
    from qsim_amd.mps import MPSChain
    case = np.load(os.path.join(HERE, 'golden', 'mps_core_samples_10q.npz'))
    n = int(case['n'])
    ch = MPSChain(n, int(case['chi_max']), float(case['cutoff']))
    for g, i in zip(case['gates'], case['sites']):
        ch.apply_2q(int(i), g)
    ch.check_status()
    for name in ('full', 'sub_a', 'sub_b', 'sub_c'):
        meas = [int(q) for q in case[f'{name}_measured']] or None
        exp = [str(s) for s in case[f'{name}_samples']]
        got = ch.sample(len(exp), meas, np.random.default_rng(int(case[f'{name}_rng_seed'])))
        assert got == exp, name


def test_dropin_run_equals_reference_compiler_fixture(sim):
    """FPGAMPSSimulator.run on a circuit with Toffolis whose adjacent sub-CX gates have
    the control site above the target, a routed gate, a reset, an unknown gate name and
    a measured subset: statevector, bond dims and seeded counts equal what the
    reference's compile_circuit + gate loop + sampling produce
    (tests/golden/mps_compile_ccx_6q.npz)."""
    from collections import Counter
    from golden.fixture_circuits import ccx_circuit
    case = np.load(os.path.join(HERE, 'golden', 'mps_compile_ccx_6q.npz'))
    qc = ccx_circuit()
    res = sim.run(qc, shots=int(case['shots']), seed=int(case['rng_seed']))
    sv = res.get_statevector()
    ref = case['statevector']
    assert np.max(np.abs(sv - ref)) < 1e-10 * max(1.0, float(np.max(np.abs(ref))))
    assert res.get_bond_dims() == [int(b) for b in case['bond_dims']]
    assert res.get_counts() == dict(Counter(str(s) for s in case['samples']))


# ------------------------------------------------ one-site gates and resets on the device
def test_batched_one_site_gates_and_device_reset_match_oracle():
    """qs_mps_apply_1q_batch over more sites than one launch carries (32) and
    qs_mps_reset_site, including the |0> fallback of a site polarised into |1>
    (fpga_mps_simulator.py:530-552), against the oracle's gate loop."""
    from qsim_amd.mps import MPSChain
    n, chi = 40, 16
    rng = np.random.default_rng(8)
    ch = MPSChain(n, chi, 1e-12)
    gl = [{'type': '2q', 'sites': [i, i + 1], 'mat': g}
          for _, i, g in workloads.brickwall_gates(n, 4, seed=8, kind='haar')]
    us = [mps_ref.gate_1q('u3', list(rng.uniform(0, 6.28, 3))) for _ in range(n)]
    for g in gl:
        ch.apply_2q(g['sites'][0], g['mat'])
    ch.apply_1q_batch([(s, us[s]) for s in range(n)])
    ch.reset_site(5)
    ch.reset_site(17)
    ch.check_status()
    ref_gl = gl + [{'type': '1q', 'sites': [s], 'mat': us[s]} for s in range(n)] + \
        [{'type': 'reset', 'sites': [5]}, {'type': 'reset', 'sites': [17]}]
    ref = mps_ref.run_gate_list(n, ref_gl, chi, 1e-12)
    got = as_oracle(ch.to_host())
    assert got.bond_dims() == ref.bond_dims()
    gg, rr = mps_ref.overlap(got, got).real, mps_ref.overlap(ref, ref).real
    gr = mps_ref.overlap(got, ref)
    assert abs(gg / rr - 1) < 1e-10                 # same norm: the resets scale alike
    assert abs(gr / rr - 1) < 1e-10                 # same state, phase included
    # a product-state site entirely in |1>: reset falls back to |0>
    one = MPSChain(3, 4, 1e-12)
    one.apply_1q_batch([(1, np.array([[0, 1], [1, 0]], dtype=complex))])
    one.reset_site(1)
    t = one.to_host().sites[1]
    assert t.shape == (1, 2, 1) and t[0, 0, 0] == 1.0 and t[0, 1, 0] == 0.0
