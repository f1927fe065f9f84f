"""Per-plan pass kernels (sv_jit.cpp) on the host: the generated source and its hipRTC
compile for gfx950, without a device.  Their results on the GPU are compared with the
op-stream interpreter bit for bit in test_sv_jit_gpu.py."""

import math
import re

import numpy as np
import pytest

from qsim_amd import workloads
from qsim_amd.sv import gates_to_records, jit_check, jit_source, plan_json
from test_sv_plan_cpu import random_circuit


def _ops(pj):
    return [[G['kind'] for S in P['stages'] for G in S['gates']] for P in pj['passes']]


def test_source_runs_the_plan_ops_in_order():
    """One op per op of the pass (marked in the source), with the plan's kinds in the plan's
    order."""
    rng = np.random.default_rng(21)
    nq = 14
    recs, _ = gates_to_records(random_circuit(rng, nq, 250), nq)
    kinds = _ops(plan_json(nq, recs))
    assert len(kinds) >= 2
    for p, want in enumerate(kinds):
        src = jit_source(nq, recs, p)
        got = [int(k) for k in re.findall(r'// op \d+ kind (\d+)', src)]
        assert got == want
        idx = [int(k) for k in re.findall(r'// op (\d+) kind', src)]
        assert idx == list(range(len(want)))
        assert 'qs_sv_pass_0' in src and 'pass_load<R>' in src and 'pass_store<R>' in src


def test_source_does_not_depend_on_angles():
    """Rotation parameters stay in the op table: the same circuit structure with other
    angles gives the same kernel text (so compiled kernels are reused)."""
    nq = 13
    a, _ = gates_to_records(workloads.u3_cz_layers(nq, 3, seed=1), nq)
    b, _ = gates_to_records(workloads.u3_cz_layers(nq, 3, seed=2), nq)
    pa, pb = plan_json(nq, a), plan_json(nq, b)
    if _ops(pa) != _ops(pb):
        pytest.skip('the planner chose different op kinds for the two angle sets')
    for p in range(len(pa['passes'])):
        assert jit_source(nq, a, p) == jit_source(nq, b, p)


def test_random_plan_compiles_for_gfx950():
    """Every op kind the planner emits on a random circuit (controlled and two-qubit
    gates, diagonals, phase groups, swaps, layouts) compiles with hipRTC."""
    rng = np.random.default_rng(5)
    nq = 12
    recs, _ = gates_to_records(random_circuit(rng, nq, 220), nq)
    kinds = {k for ops in _ops(plan_json(nq, recs)) for k in ops}
    assert {6, 10}.issubset(kinds) and len(kinds) >= 6
    nk, nbytes = jit_check(nq, recs)
    assert 1 <= nk <= len(plan_json(nq, recs)['passes'])
    assert nbytes > 1000


def test_narrow_state_compiles():
    """States below a full tile use narrow register blocks (R < 4)."""
    for nq in (1, 3, 6):
        recs, _ = gates_to_records([('h', [0], None), ('rz', [0], [0.3])] +
                                   ([('cx', [0, nq - 1], None)] if nq > 1 else []), nq)
        nk, nbytes = jit_check(nq, recs)
        assert nk >= 1 and nbytes > 0
