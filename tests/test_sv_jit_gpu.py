"""Per-plan pass kernels (qs_sv_plan_specialize) against the op-stream interpreter on the
GPU: the same ops in the same order, so the states must agree bit for bit.  The
interpreter itself is checked against the SV oracle in test_sv_gpu.py."""

import numpy as np
import pytest

from qsim_amd import workloads
from test_sv_plan_cpu import random_circuit

pytestmark = pytest.mark.gpu


def _pair(nq, recs):
    import torch
    from qsim_amd.sv import StateVector
    ref = StateVector(nq, 0)
    got = StateVector(nq, 0)
    plan = got.plan(recs)
    plan.specialize()
    ref.apply_records(recs)
    got.launch(plan)
    torch.cuda.synchronize()
    return ref, got, plan


@pytest.mark.parametrize('nq,ng,seed', [(1, 20, 0), (4, 120, 1), (9, 250, 2), (12, 300, 3),
                                        (14, 300, 4), (16, 300, 5), (24, 300, 9)])
def test_random_circuit_specialized_equals_interpreter(nq, ng, seed):
    import torch
    from qsim_amd.sv import gates_to_records
    print(f'specialized random circuit nq={nq} ng={ng} seed={seed}', flush=True)
    recs, _ = gates_to_records(random_circuit(np.random.default_rng(seed), nq, ng), nq)
    ref, got, _ = _pair(nq, recs)
    assert torch.equal(ref.tensor, got.tensor)


@pytest.mark.parametrize('nq', [14, 20])
def test_benchmark_structure_specialized_equals_interpreter(nq):
    """U3 layers with CZ ladders and the QFT (the 28-qubit workload's structure), twice:
    the second launch reuses the plan's kernels and its block-factor table."""
    import torch
    from qsim_amd.sv import gates_to_records
    print(f'specialized benchmark structure nq={nq}', flush=True)
    recs, _ = gates_to_records(workloads.sv_benchmark_circuit(nq, 6, 9), nq)
    ref, got, plan = _pair(nq, recs)
    assert torch.equal(ref.tensor, got.tensor)
    ref.apply_records(recs)
    got.launch(plan)
    torch.cuda.synchronize()
    assert torch.equal(ref.tensor, got.tensor)


def test_kernel_cache_reused_across_angles_and_graph_replay():
    """A plan of the same structure with other angles compiles nothing new, and the
    specialised launch replays from a HIP graph (three replays with eager work and
    synchronisations between them) with the interpreter's result each time."""
    import torch
    from qsim_amd.sv import StateVector, gates_to_records, jit_source
    nq = 15
    a, _ = gates_to_records(workloads.u3_cz_layers(nq, 4, seed=11), nq)
    b, _ = gates_to_records(workloads.u3_cz_layers(nq, 4, seed=12), nq)
    ref = StateVector(nq, 0)
    got = StateVector(nq, 0)
    pa = got.plan(a)
    pa.specialize()
    pb = got.plan(b)
    same = pa.npasses == pb.npasses and all(
        jit_source(nq, a, p) == jit_source(nq, b, p) for p in range(pa.npasses))
    compiled = pb.specialize()
    if same:
        assert compiled == 0
    got.launch(pb)
    ref.apply_records(b)
    torch.cuda.synchronize()
    assert torch.equal(ref.tensor, got.tensor)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        got.launch(pb)
    torch.cuda.current_stream().wait_stream(s)
    ref.apply_records(b)
    torch.cuda.synchronize()
    assert torch.equal(ref.tensor, got.tensor)
    with torch.cuda.graph(g):
        got.launch(pb)
    for _ in range(3):
        ref.apply_records(b)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ref.tensor, got.tensor)
