"""GPU parity tests of the statevector path (libqsim_hip.so) against the SV oracle.

Known-answer tests follow comprehensive_test_v07.py (sections 1-4, 5, 6, 8) and the
MCX stress table (mcx_stress_results.csv).  Random circuits are compared amplitude by
amplitude with the oracle (reference lowering + pair updates) at 1e-12.  At 28 qubits,
where the oracle is too slow, size-independent properties are checked: norm, a
circuit followed by its inverse, and bitwise determinism.
"""

import csv
import math
import os

import numpy as np
import pytest

from oracle import sv_ref
from qsim_amd import workloads
from test_sv_plan_cpu import random_circuit

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def fpga():
    from fpga_simulator_v07 import QuantumFPGACircuit
    return QuantumFPGACircuit('unused.xclbin')


def run(fpga, nq, gates, **kw):
    res = fpga.run_circuit(nq, gates, shots=kw.pop('shots', 0), **kw)
    return res


def prob(sv, i):
    return float(abs(sv[i]) ** 2)


# ---------------------------------------------------------------- section 1-4 KATs
KAT_DET = [
    (3, [('id', [0], None)], 0),
    (1, [('x', [0], None)], 1),
    (1, [('x', [0], None), ('x', [0], None)], 0),
    (1, [('y', [0], None)], 1),
    (1, [('z', [0], None)], 0),
    (1, [('s', [0], None), ('sdg', [0], None)], 0),
    (1, [('t', [0], None), ('tdg', [0], None)], 0),
    (1, [('sx', [0], None), ('sxdg', [0], None)], 0),
    (1, [('sx', [0], None), ('sx', [0], None)], 1),
    (1, [('rx', [0], [math.pi])], 1),
    (1, [('rx', [0], [2 * math.pi])], 0),
    (1, [('ry', [0], [math.pi])], 1),
    (1, [('rz', [0], [math.pi])], 0),
    (1, [('p', [0], [math.pi])], 0),
    (1, [('u3', [0], [math.pi, 0.0, math.pi])], 1),
    (2, [('x', [0], None), ('cx', [0, 1], None)], 3),
    (2, [('x', [0], None), ('cx', [0, 1], None), ('cx', [0, 1], None), ('x', [0], None)], 0),
    (2, [('cz', [0, 1], None)], 0),
    (2, [('x', [0], None), ('x', [1], None), ('cz', [0, 1], None)], 3),
    (2, [('x', [0], None), ('cy', [0, 1], None)], 3),
    (2, [('x', [0], None), ('swap', [0, 1], None)], 2),
    (2, [('x', [0], None), ('swap', [0, 1], None), ('swap', [0, 1], None), ('x', [0], None)], 0),
    (2, [('x', [0], None), ('x', [1], None), ('cp', [0, 1], [math.pi])], 3),
    (2, [('x', [0], None), ('x', [1], None), ('cp', [0, 1], [2 * math.pi]),
         ('x', [0], None), ('x', [1], None)], 0),
    (2, [('x', [0], None), ('crx', [0, 1], [math.pi])], 3),
    (2, [('x', [0], None), ('cry', [0, 1], [math.pi])], 3),
    (2, [('x', [0], None), ('csx', [0, 1], None), ('csx', [0, 1], None)], 3),
    (2, [('x', [0], None), ('iswap', [0, 1], None)], 2),
    (3, [('x', [0], None), ('x', [1], None), ('ccx', [0, 1, 2], None)], 7),
    (3, [('x', [0], None), ('ccx', [0, 1, 2], None)], 1),
    (3, [('x', [0], None), ('x', [1], None), ('ccx', [0, 1, 2], None), ('ccx', [0, 1, 2], None),
         ('x', [0], None), ('x', [1], None)], 0),
]


@pytest.mark.parametrize('nq,gates,det', KAT_DET)
def test_kat_deterministic(fpga, nq, gates, det):
    res = run(fpga, nq, gates, det_idx=det, det_tolerance=1e-12)
    assert res['validation'].passed, res['validation']
    assert prob(res['statevector'], det) > 1 - 1e-12
    ref = sv_ref.run_circuit(nq, gates)
    assert np.max(np.abs(res['statevector'] - ref)) < 1e-12


@pytest.mark.parametrize('name,params', [('ch', None), ('crz', [math.pi]), ('csx', None),
                                         ('dcx', None), ('iswap', None), ('ecr', None),
                                         ('rxx', [math.pi / 2]), ('ryy', [math.pi / 2]),
                                         ('rzz', [math.pi / 2])])
def test_kat_norm_and_amplitudes(fpga, name, params):
    gates = [('x', [0], None), (name, [0, 1], params)]
    res = run(fpga, 2, gates)
    assert res['validation'].passed
    assert np.max(np.abs(res['statevector'] - sv_ref.run_circuit(2, gates))) < 1e-12


@pytest.mark.parametrize('prep,probs', [([], {1: .5, 3: .5}), ([('x', [0], None)], {0: .5, 2: .5}),
                                        ([('x', [1], None)], {1: .5, 3: .5}),
                                        ([('x', [0], None), ('x', [1], None)], {0: .5, 2: .5})])
def test_ecr_columns(fpga, prep, probs):
    res = run(fpga, 2, prep + [('ecr', [0, 1], None)])
    for i, p in probs.items():
        assert abs(prob(res['statevector'], i) - p) < 1e-12


def test_uniform_h_u2(fpga):
    for gates in ([('h', [0], None)], [('u2', [0], [0.0, math.pi])],
                  [('u3', [0], [math.pi / 2, 0.0, math.pi])]):
        res = run(fpga, 1, gates, uniform_check=True, uniform_tolerance=1e-12)
        assert res['validation'].passed


# ---------------------------------------------------------------- section 5: QFT
@pytest.mark.parametrize('nq', [1, 2, 3, 5, 8, 11, 12, 13, 14, 16, 20])
def test_qft_uniform_and_matches_oracle(fpga, nq):
    gates = workloads.qft_gates(nq)
    res = run(fpga, nq, gates, uniform_check=True, uniform_tolerance=1e-10)
    assert res['validation'].passed
    if nq <= 16:
        ref = sv_ref.run_circuit(nq, gates)
        assert np.max(np.abs(res['statevector'] - ref)) < 1e-12


@pytest.mark.parametrize('nq', [1, 2, 4, 8, 12, 15])
def test_qft_iqft_roundtrip(fpga, nq):
    gates = workloads.qft_gates(nq) + workloads.qft_gates(nq, inverse=True)
    res = run(fpga, nq, gates, det_idx=0, det_tolerance=1e-10)
    assert res['validation'].passed


# ---------------------------------------------------------------- section 6: Grover
def grover(n):
    def mcz(qs):
        if len(qs) == 1:
            return [('z', [qs[0]], None)]
        if len(qs) == 2:
            return [('cz', qs, None)]
        if len(qs) == 3:
            return [('h', [qs[2]], None), ('ccx', qs, None), ('h', [qs[2]], None)]
        return [('h', [qs[-1]], None), ('mcx', qs, None), ('h', [qs[-1]], None)]
    th = math.asin(1 / math.sqrt(1 << n))
    iters = max(1, round(math.pi / (4 * th) - 0.5))
    g = [('h', [q], None) for q in range(n)]
    for _ in range(iters):
        g += mcz(list(range(n)))
        g += [('h', [q], None) for q in range(n)] + [('x', [q], None) for q in range(n)]
        g += mcz(list(range(n)))
        g += [('x', [q], None) for q in range(n)] + [('h', [q], None) for q in range(n)]
    return g


@pytest.mark.parametrize('n', [2, 3, 4, 5, 6, 8, 10])
def test_grover_marks_all_ones(fpga, n):
    g = grover(n)
    res = run(fpga, n, g)
    p = prob(res['statevector'], (1 << n) - 1)
    assert p >= 0.5
    ref = sv_ref.run_circuit(n, g)
    assert abs(p - abs(ref[(1 << n) - 1]) ** 2) < 1e-12


# ---------------------------------------------------------------- MCX stress table
def test_mcx_stress_table(fpga):
    with open(os.path.join(HERE, 'golden', 'mcx_stress_expected.csv')) as f:
        rows = list(csv.DictReader(f))
    seen = 0
    for r in rows:
        n = int(r['n_qubits'])
        ctrls, tgt = list(range(n - 1)), n - 1
        bits = [int(c) for c in r['pattern_bits']] if r['pattern_bits'] != '-' else []
        kind = r['test_type']
        prep = [('x', [c], None) for c, b in zip(ctrls, bits) if b]
        if kind in ('FIRE', 'NO-FIRE'):
            g = prep + [('mcx', ctrls + [tgt], None)]
        elif kind == 'DOUBLE':
            g = prep + [('mcx', ctrls + [tgt], None)] * 2
        else:  # PHASE
            g = prep + [('h', [tgt], None), ('mcx', ctrls + [tgt], None), ('h', [tgt], None)]
        res = run(fpga, n, g)
        assert prob(res['statevector'], int(r['expected_idx'])) > 1 - 1e-12, r
        seen += 1
    assert seen == len(rows) > 500


# ---------------------------------------------------------------- section 8: clbits
def test_measurement_clbit_mapping(fpga):
    np.random.seed(7)
    r = fpga.run_circuit(1, [('h', [0], None), ('measure', [0], [0]), ('measure', [0], [1])],
                         shots=2000, num_clbits=3)
    assert set(r['counts']) <= {'000', '011'} and len(r['counts']) == 2
    r = fpga.run_circuit(1, [('x', [0], None), ('measure', [0], [2])], shots=200, num_clbits=3)
    assert set(r['counts']) == {'100'}
    r = fpga.run_circuit(2, [('h', [0], None), ('x', [1], None), ('measure', [0], [0]),
                             ('measure', [1], [1])], shots=500, num_clbits=2)
    assert set(r['counts']) <= {'10', '11'}
    r = fpga.run_circuit(2, [('x', [0], None), ('measure', [0, 1], None)], shots=100,
                         num_clbits=2)
    assert set(r['counts']) == {'01'}
    r = fpga.run_circuit(2, [('id', [0], None)], shots=100, num_clbits=3)
    assert set(r['counts']) == {'000'}


def _host_choice(sv, u):
    # np.random.choice(N, len(u), p=probs) for the uniforms u it would draw
    # (numpy's legacy choice: cdf = cumsum(p), cdf /= cdf[-1], searchsorted right)
    p = sv.real ** 2 + sv.imag ** 2
    p = p / p.sum()
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    return np.searchsorted(cdf, u, side='right')


@pytest.mark.parametrize('nq,kind', [(1, 'h'), (10, 'random'), (14, 'ghz'), (18, 'random'),
                                     (20, 'sparse')])
def test_device_sampler_matches_numpy_choice(nq, kind):
    from qsim_amd.sv import StateVector, gates_to_records
    rng = np.random.default_rng(nq)
    if kind == 'h':
        gates = [('h', [0], None)]
    elif kind == 'ghz':
        gates = [('h', [0], None)] + [('cx', [q, q + 1], None) for q in range(nq - 1)]
    elif kind == 'sparse':
        # probability only on the basis states whose top qubits are 0: many whole blocks
        # and lanes of zero probability
        gates = [('h', [q], None) for q in range(6)] + [('ry', [q], [0.3 * q]) for q in range(6)]
    else:
        gates = random_circuit(rng, nq, 8 * nq)
    sv = StateVector(nq, 0)
    sv.apply_records(gates_to_records(gates, nq)[0])
    host = sv.to_numpy()
    u = rng.random(5000)
    u[0] = 0.0                       # the first basis state of nonzero probability
    got = sv.sample_indices(u)
    assert np.array_equal(got, _host_choice(host, u))
    assert np.all(np.abs(host[got]) > 0)
    if kind == 'ghz':
        assert set(got.tolist()) <= {0, (1 << nq) - 1}


def test_drop_in_shots_follow_the_global_generator(fpga):
    # the reference's run_circuit samples with np.random.choice on the global generator;
    # the drop-in draws the same uniforms and samples on the device
    g = [('h', [0], None), ('ry', [1], [0.7]), ('cx', [1, 2], None), ('rx', [3], [1.1])]
    np.random.seed(123)
    r = fpga.run_circuit(4, g, shots=3000)
    np.random.seed(123)
    p = np.abs(r['statevector']) ** 2
    idx = np.random.choice(16, size=3000, p=p / p.sum())
    exp = [format(int(i), '04b') for i in idx]
    assert r['memory'] == exp


def test_reset_matches_oracle(fpga):
    g = [('h', [0], None), ('cx', [0, 1], None), ('reset', [0], None), ('h', [2], None),
         ('reset', [2], None), ('x', [2], None), ('reset', [2], None), ('ry', [1], [0.3])]
    res = run(fpga, 3, g)
    assert np.max(np.abs(res['statevector'] - sv_ref.run_circuit(3, g))) < 1e-12
    # reset of a qubit that is fully |1> falls back to |0..0>
    res = run(fpga, 1, [('x', [0], None), ('reset', [0], None)])
    assert prob(res['statevector'], 0) > 1 - 1e-15


def test_distribute_arbitrary_state_skip_init(fpga):
    rng = np.random.default_rng(3)
    nq = 9
    v = rng.standard_normal(1 << nq) + 1j * rng.standard_normal(1 << nq)
    v /= np.linalg.norm(v)
    g = random_circuit(rng, nq, 80)
    fpga._distribute_arbitrary_state(v, nq)
    res = fpga.run_circuit(nq, g, shots=0, skip_init=True)
    ref = sv_ref.run_circuit(nq, g, init=v)
    assert np.max(np.abs(res['statevector'] - ref)) < 1e-12


# ---------------------------------------------------------------- random circuits
@pytest.mark.parametrize('nq,ng,seed', [(4, 200, 10), (9, 300, 11), (12, 400, 12),
                                        (13, 400, 13), (16, 500, 14), (18, 400, 15),
                                        (20, 300, 16)])
def test_random_circuits_match_oracle(fpga, nq, ng, seed):
    rng = np.random.default_rng(seed)
    g = random_circuit(rng, nq, ng)
    res = run(fpga, nq, g)
    ref = sv_ref.run_circuit(nq, g)
    assert np.max(np.abs(res['statevector'] - ref)) < 1e-12


@pytest.mark.parametrize('nq', [14, 18])
def test_benchmark_structure_matches_oracle(fpga, nq):
    # the 28-qubit benchmark's structure at sizes the oracle finishes: fused rotation
    # runs, sign groups with thread and thread-pair terms, block-class phase factors,
    # complex phase groups and register SWAPs of the QFT
    g = workloads.u3_cz_layers(nq, 20, seed=5) + workloads.qft_gates(nq)
    res = run(fpga, nq, g)
    ref = sv_ref.run_circuit(nq, g)
    assert np.max(np.abs(res['statevector'] - ref)) < 1e-12


def test_deep_single_pass_circuit_matches_oracle(fpga):
    # 9000 rotations in one pass: the planner divides the scaled rotations' growth out
    rng = np.random.default_rng(41)
    nq = 6
    g = [('u3', [int(rng.integers(nq))], [float(v) for v in rng.uniform(0, 2 * math.pi, 3)])
         for _ in range(9000)]
    res = run(fpga, nq, g)
    ref = sv_ref.run_circuit(nq, g)
    assert np.all(np.isfinite(res['statevector']))
    assert np.max(np.abs(res['statevector'] - ref)) < 1e-10


def test_words_format_matches_oracle_f32():
    """The 8-word instruction entry point reproduces the FPGA's float32 parameters."""
    from qsim_amd.sv import StateVector
    ops_by_name = sv_ref.KERNEL_GATE_OPS
    rng = np.random.default_rng(21)
    nq = 10
    words, ops = [], []
    for _ in range(300):
        nm = ['h', 'x', 'y', 'z', 's', 't', 'sdg', 'tdg', 'sx', 'sxdg', 'rx', 'ry', 'rz', 'p',
              'u2', 'u3', 'cx', 'cz'][rng.integers(18)]
        if nm in ('cx', 'cz'):
            qs = [int(v) for v in rng.choice(nq, 2, replace=False)]
        else:
            qs = [int(rng.integers(nq))]
        ps = [float(v) for v in rng.uniform(-4, 4, 3)]
        w = np.zeros(8, dtype=np.int32)
        w[0] = ops_by_name[nm]
        w[1] = qs[0]
        w[2] = qs[1] if len(qs) > 1 else -1
        w[3] = -1
        w[4:7] = np.array(ps, dtype=np.float32).view(np.int32)
        words.append(w)
        ops.append((nm, qs, ps))
    sv = StateVector(nq, 0)
    sv.apply_words(np.concatenate(words))
    ref = np.zeros(1 << nq, dtype=np.complex128)
    ref[0] = 1
    sv_ref.apply_kernel_ops(ref, nq, ops, f32_params=True)
    assert np.max(np.abs(sv.to_numpy() - ref)) < 1e-12


# ---------------------------------------------------------------- 28 qubits
def test_28q_properties():
    """Full-size workload: norm, inverse round trip, determinism (no oracle at 28q)."""
    import torch
    from qsim_amd.sv import StateVector
    nq = 28
    g = workloads.u3_cz_layers(nq, 3, seed=9) + workloads.qft_gates(nq)
    sv = StateVector(nq, 0)
    sv.apply(g)
    assert abs(sv.norm_sq() - 1.0) < 1e-10
    snap = sv.tensor.clone()
    sv.apply(workloads.inverse_circuit(g))
    torch.cuda.synchronize()
    a0 = sv.tensor[0].item()
    assert abs(abs(a0) - 1.0) < 1e-10
    sv.set_zero()
    sv.apply(g)
    assert torch.equal(sv.tensor, snap)          # bitwise deterministic
    del snap


# ------------------------------------------------------- plan once, launch many
def test_plan_launch_equals_apply_bitwise_and_replays_in_a_graph():
    """qs_sv_plan_create + qs_sv_plan_launch give qs_sv_apply's result bit for bit, on
    every launch of the same plan, eagerly and replayed from a HIP graph (three replays,
    eager work and synchronisations between them)."""
    import torch
    from qsim_amd.sv import StateVector, gates_to_records
    nq = 16
    recs, _ = gates_to_records(workloads.u3_cz_layers(nq, 6, seed=3) + workloads.qft_gates(nq),
                               nq)
    ref = StateVector(nq, 0)
    got = StateVector(nq, 0)
    plan = got.plan(recs)
    assert plan.npasses >= 2 and plan.device_bytes > 0
    for _ in range(2):
        ref.apply_records(recs)
        got.launch(plan)
        torch.cuda.synchronize()
        assert torch.equal(ref.tensor, got.tensor)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        got.launch(plan)                 # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    ref.apply_records(recs)
    torch.cuda.synchronize()
    assert torch.equal(ref.tensor, got.tensor)
    with torch.cuda.graph(g):
        got.launch(plan)
    for _ in range(3):
        ref.apply_records(recs)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ref.tensor, got.tensor)
    plan.destroy()


def test_28q_depth20_benchmark_circuit_inverse_roundtrip():
    """The full 28-qubit benchmark circuit (20 x [U3 + CZ ladder] + QFT-28, 1520 gates)
    through a prebuilt plan run by its per-plan kernels (the benchmark's path): equal to the
    op-stream interpreter's run of the same plan, then its inverse: back to |0..0> and
    unit norm."""
    import torch
    from qsim_amd.sv import StateVector, gates_to_records
    nq = 28
    g = workloads.sv_benchmark_circuit(nq, 20, 1234)
    recs, _ = gates_to_records(g, nq)
    inv, _ = gates_to_records(workloads.inverse_circuit(g), nq)
    sv = StateVector(nq, 0)
    fwd = sv.plan(recs)
    fwd.specialize()
    sv.launch(fwd)
    ref = StateVector(nq, 0)
    ref_plan = ref.plan(recs)
    ref.launch(ref_plan)
    torch.cuda.synchronize()
    assert torch.equal(sv.tensor, ref.tensor)
    del ref
    ref_plan.destroy()
    assert abs(sv.norm_sq() - 1.0) < 1e-10
    sv.apply_records(inv)
    torch.cuda.synchronize()
    assert abs(abs(sv.tensor[0].item()) - 1.0) < 1e-9
    assert abs(sv.norm_sq() - 1.0) < 1e-10
    fwd.destroy()


def test_32q_indices_beyond_2_pow_31():
    """A 64 GiB state (32 qubits): amplitude indices and byte offsets past 2^31 land where
    the basis-state index says, and a layered circuit followed by its inverse returns to
    |0..0>."""
    import torch
    from qsim_amd.sv import StateVector
    nq = 32
    sv = StateVector(nq, 0)
    sv.apply([('h', [31], None), ('cx', [31, 0], None), ('x', [30], None)])
    r = 1 / math.sqrt(2)
    i0, i1 = 1 << 30, (1 << 31) | (1 << 30) | 1
    assert abs(sv.tensor[i0].item() - r) < 1e-15
    assert abs(sv.tensor[i1].item() - r) < 1e-15
    assert abs(sv.norm_sq() - 1.0) < 1e-12
    sv.set_zero()
    g = workloads.u3_cz_layers(nq, 2, seed=13)
    sv.apply(g)
    sv.apply(workloads.inverse_circuit(g))
    torch.cuda.synchronize()
    assert abs(abs(sv.tensor[0].item()) - 1.0) < 1e-10
    del sv
    torch.cuda.empty_cache()
