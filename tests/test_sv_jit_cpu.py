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


def _layout(lit):
    g = [list(map(int, x.split(','))) for x in re.findall(r'\{([-\d,]*)\}', lit)]
    return {'regs': g[0][:4], 'tpos': g[2]}


def _split_moves(fr, to, p, nthr=256, nreg=16, half=2176):
    """Model of layout_split's addressing (sv_ops.h): per round h, the LDS slot every
    (thread, register) pair writes under `fr` and reads under `to`.  Returns the
    (thread, register) -> local index map the reads produce."""
    def slot(l):
        return l + (l >> 4)

    def drop(l):
        return (l & ((1 << p) - 1)) | ((l >> (p + 1)) << p)

    def loc_t(D, t):
        return sum(((t >> b) & 1) << D['tpos'][b] for b in range(len(D['tpos'])))

    def rl(D, j):
        return sum(1 << D['regs'][i] for i in range(4) if (j >> i) & 1)

    fi = fr['regs'].index(p) if p in fr['regs'] else -1
    ti = to['regs'].index(p) if p in to['regs'] else -1
    fb = -1 if fi >= 0 else fr['tpos'].index(p)
    tb = -1 if ti >= 0 else to['tpos'].index(p)
    got = {}
    for h in range(2):
        lds = {}
        for t in range(nthr):
            base = slot(drop(loc_t(fr, t) & ~(1 << p)))
            for j in range(nreg):
                if (fi >= 0 and ((j >> fi) & 1) != h) or (fi < 0 and ((t >> fb) & 1) != h):
                    continue
                s = base + slot(drop(rl(fr, j) & ~(1 << p)))
                assert 0 <= s < half and s not in lds
                lds[s] = loc_t(fr, t) | rl(fr, j)
        for t in range(nthr):
            base = slot(drop(loc_t(to, t) & ~(1 << p)))
            for j in range(nreg):
                if (ti >= 0 and ((j >> ti) & 1) != h) or (ti < 0 and ((t >> tb) & 1) != h):
                    continue
                got[(t, j)] = lds[base + slot(drop(rl(to, j) & ~(1 << p)))]
    return got


def test_split_transposes_move_every_amplitude():
    """Each half-tile transpose the generator emits (split position included) writes every
    amplitude once into 2176 LDS slots per round and reads back, for each (thread,
    register) of the new layout, the amplitude with that local index (a model of
    layout_split's addressing over three passes of the benchmark plan)."""
    nq = 28
    recs, _ = gates_to_records(workloads.sv_benchmark_circuit(nq, 20, 1234), nq)
    pat = re.compile(r'constexpr DevLayout from = (\{\{.*?\}\}); constexpr DevLayout to = '
                     r'(\{\{.*?\}\}); layout_split<R>\(a, Lcur, from, to, (\d+), lds, t\)')
    n = 0
    npass = len(plan_json(nq, recs)['passes'])
    for p in (0, 1, npass - 1):
        for fr, to, pos in pat.findall(jit_source(nq, recs, p)):
            fr, to = _layout(fr), _layout(to)
            got = _split_moves(fr, to, int(pos))
            for (t, j), v in got.items():
                want = sum(((t >> b) & 1) << to['tpos'][b] for b in range(8)) | \
                    sum(1 << to['regs'][i] for i in range(4) if (j >> i) & 1)
                assert v == want
            assert len(got) == 4096
            n += 1
    assert n >= 10
