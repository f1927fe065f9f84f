"""Host-side checks of the statevector path (no GPU): gate records and the pass planner.

The plan produced by libqsim_hip.so's planner is replayed with NumPy the way the
pass kernel reads it and compared with the SV oracle (reference lowering + pair
updates), including the global phase the reference lowering produces.
"""

import math

import numpy as np
import pytest

from oracle import sv_ref
from qsim_amd import _lib
from qsim_amd.sv import gates_to_records, plan_json
from sv_plan_replay import replay

NAMES_1Q = ['h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx', 'sxdg', 'rx', 'ry', 'rz', 'p',
            'u1', 'u2', 'u3', 'u', 'phase', 'id']
NAMES_2Q = ['cx', 'cnot', 'cz', 'cy', 'ch', 'swap', 'cp', 'crx', 'cry', 'crz', 'csx', 'dcx',
            'iswap', 'ecr', 'rxx', 'ryy', 'rzz']
NPARAM = {'rx': 1, 'ry': 1, 'rz': 1, 'p': 1, 'u1': 1, 'phase': 1, 'u2': 2, 'u3': 3, 'u': 3,
          'cp': 1, 'crx': 1, 'cry': 1, 'crz': 1, 'rxx': 1, 'ryy': 1, 'rzz': 1}


def random_circuit(rng, nq, ng, with_ccx=True):
    gates = []
    for _ in range(ng):
        r = rng.random()
        if r < 0.45 or nq < 2:
            nm = NAMES_1Q[rng.integers(len(NAMES_1Q))]
            qs = [int(rng.integers(nq))]
        elif r < 0.93 or nq < 3 or not with_ccx:
            nm = NAMES_2Q[rng.integers(len(NAMES_2Q))]
            qs = [int(v) for v in rng.choice(nq, 2, replace=False)]
        else:
            nm = 'ccx'
            qs = [int(v) for v in rng.choice(nq, 3, replace=False)]
        ps = [float(v) for v in rng.uniform(-math.pi, math.pi, NPARAM.get(nm, 0))] or None
        gates.append((nm, qs, ps))
    return gates


def _lib_or_skip():
    try:
        _lib.load_library()
    except Exception as e:  # the .so must be built by __graft_entry__.build()
        pytest.fail(f'libqsim_hip.so not loadable: {e}')


def run_plan_cpu(nq, gates):
    recs, phase = gates_to_records(gates, nq)
    plan = plan_json(nq, recs)
    st = np.zeros(1 << nq, dtype=np.complex128)
    st[0] = 1
    out = replay(plan, nq, st) * phase
    return out / np.linalg.norm(out), plan


@pytest.mark.parametrize('nq,ng,seed', [(1, 10, 0), (2, 30, 1), (3, 60, 2), (5, 120, 3),
                                        (12, 150, 4), (13, 200, 5), (14, 250, 7),
                                        (15, 300, 6), (16, 300, 8)])
def test_plan_replay_matches_oracle(nq, ng, seed):
    _lib_or_skip()
    rng = np.random.default_rng(seed)
    gates = random_circuit(rng, nq, ng)
    got, plan = run_plan_cpu(nq, gates)
    ref = sv_ref.run_circuit(nq, gates)
    assert np.max(np.abs(got - ref)) < 1e-12
    # tiles of 12 qubits (4096 amplitudes) once the state has more
    assert all(p['K'] == min(nq, 12) for p in plan['passes'])


def test_plan_replay_benchmark_structure():
    # 20 x [U3 on every qubit + CZ ladder] + QFT at 14 qubits: fused rotation runs, sign
    # and complex phase groups, block-class factors and register SWAPs, replayed as the
    # kernel reads them
    _lib_or_skip()
    from qsim_amd import workloads
    nq = 14
    gates = workloads.u3_cz_layers(nq, 20, seed=5) + workloads.qft_gates(nq)
    got, plan = run_plan_cpu(nq, gates)
    ref = sv_ref.run_circuit(nq, gates)
    assert np.max(np.abs(got - ref)) < 1e-12
    kinds = {g['kind'] for p in plan['passes'] for s in p['stages'] for g in s['gates']}
    assert 9 in kinds and 7 in kinds and 5 in kinds          # SWAP, sign and complex groups
    assert any(k >= 32 for k in kinds)                        # fused rotation runs


@pytest.mark.parametrize('nq', [13, 15, 18, 24])
def test_tile_index_map_is_a_bijection(nq):
    # Each pass's run tables send the tile index bits one to one onto the qubits outside
    # the tile; with more than three of them, index bits 0..2 (which pick the XCD a
    # workgroup runs on) land on the three highest, the rest follow in ascending order.
    _lib_or_skip()
    from qsim_amd import workloads
    gates = workloads.u3_cz_layers(nq, 4, seed=nq) + workloads.qft_gates(nq)
    recs, _ = gates_to_records(gates, nq)
    plan = plan_json(nq, recs)
    for p in plan['passes']:
        block = [q for q in range(nq) if q not in p['tq']]
        bits = {}
        for src, dst, mask in p['runs']:
            assert mask & (mask + 1) == 0, 'a run covers consecutive index bits'
            for i in range(mask.bit_length()):
                assert src + i not in bits
                bits[src + i] = dst + i
        m = [bits[i] for i in range(len(bits))]
        assert sorted(m) == block
        if len(block) > 3:
            assert m == block[-3:] + block[:-3]
        else:
            assert m == block


def test_plan_u3_cz_layers_pass_count():
    _lib_or_skip()
    nq = 20
    gates = sv_ref.u3_cz_layers(nq, 4, seed=3)
    got, plan = run_plan_cpu(nq, gates)
    ref = sv_ref.run_circuit(nq, gates)
    assert np.max(np.abs(got - ref)) < 1e-12
    # 16 non-low qubits per layer, 8 new per pass: about 2 passes per layer
    assert len(plan['passes']) <= 2 * 4 + 2


def test_benchmark_plan_takes_light_cone_windows():
    """The 28-qubit benchmark circuit (20 x [U3 on every qubit + CZ ladder] + QFT-28): the
    tile choice advances light cones over windows of consecutive qubits, 17 passes instead
    of the 31 that first-come tiles take, and every one-qubit rotation lands in exactly one
    pass, on a tile qubit."""
    _lib_or_skip()
    from qsim_amd import workloads
    nq = 28
    recs, _ = gates_to_records(workloads.sv_benchmark_circuit(nq, 20, 1234), nq)
    plan = plan_json(nq, recs)
    assert len(plan['passes']) <= 17
    nrot = 0
    for p in plan['passes']:
        assert p['K'] == 12 and p['tq'][:4] == [0, 1, 2, 3] and len(set(p['tq'])) == 12
        for s in p['stages']:
            for g in s['gates']:
                k = g['kind']
                nrot += 1 if 16 <= k < 32 else (bin(k - 31).count('1') if k >= 32 else 0)
    assert nrot == 20 * nq + nq           # the U3 layers and the QFT's Hadamards


def test_plan_qft_roundtrip_and_phase():
    _lib_or_skip()
    nq = 14
    gates = sv_ref.qft_gates(nq)
    got, _ = run_plan_cpu(nq, gates)
    ref = sv_ref.run_circuit(nq, gates)
    # QFT|0> is uniform; the reference lowering adds exp(i*theta/4) per cp
    assert np.allclose(np.abs(ref), 1 / math.sqrt(2 ** nq), atol=1e-12)
    assert np.max(np.abs(got - ref)) < 1e-12
    got2, _ = run_plan_cpu(nq, sv_ref.qft_gates(nq) + sv_ref.qft_gates(nq, inverse=True))
    assert abs(abs(got2[0]) - 1) < 1e-12


def test_mcx_records_match_permutation():
    _lib_or_skip()
    nq = 6
    for pattern in range(1 << 5):
        gates = [('x', [c], None) for c in range(5) if (pattern >> c) & 1]
        gates.append(('mcx', [0, 1, 2, 3, 4, 5], None))
        got, _ = run_plan_cpu(nq, gates)
        exp = pattern | ((1 << 5) if pattern == 31 else 0)
        assert abs(abs(got[exp]) - 1) < 1e-12


def test_bad_gates_raise():
    _lib_or_skip()
    with pytest.raises(ValueError):
        gates_to_records([('bogus', [0], None)], 2)
    with pytest.raises(ValueError):
        gates_to_records([('cx', [0, 0], None)], 2)
    with pytest.raises(ValueError):
        gates_to_records([('h', [5], None)], 2)
    # the library rejects what slips past the Python layer
    recs, _ = gates_to_records([('h', [0], None)], 2)
    bad = [(recs[0][0], 7, 0, 0, recs[0][4])]
    with pytest.raises(_lib.QsError):
        plan_json(2, bad)


def test_deep_single_pass_circuit_stays_finite():
    """Thousands of rotations in one pass (a 6-qubit state is one tile): the scaled
    rotations' growth is divided out along the way, so nothing overflows."""
    _lib_or_skip()
    rng = np.random.default_rng(41)
    nq = 6
    gates = [('u3', [int(rng.integers(nq))], [float(v) for v in rng.uniform(0, 2 * math.pi, 3)])
             for _ in range(9000)]
    got, plan = run_plan_cpu(nq, gates)
    assert len(plan['passes']) == 1
    ref = sv_ref.run_circuit(nq, gates)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) < 1e-10
