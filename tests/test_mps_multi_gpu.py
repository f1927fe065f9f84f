"""Batched independent circuits (uneven n, chi_max, depth) against the MPS oracle."""

import numpy as np
import pytest

from oracle import mps_ref

pytestmark = pytest.mark.gpu


def oracle_run(spec, cutoff):
    from qsim_amd.mps_multi import near_identity_batch
    rng = np.random.default_rng(spec.seed)
    m = mps_ref.MPS(spec.n, spec.chi_max, cutoff)
    for r in range(spec.depth):
        sites = list(range(r % 2, spec.n - 1, 2))
        if not sites:
            continue
        g = near_identity_batch(rng, len(sites))
        for i, gi in zip(sites, g):
            m.apply_adjacent_2q(i, gi)
    return m


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_multi_circuit_batch_matches_oracle(precision):
    from qsim_amd.mps_multi import CircuitSpec, MultiRun
    # the chi-100 and chi-128 circuits reach their chi_max inside the batch (bonds 100
    # and 128 in the middle), so both SVD classes and, for fp32, both MFMA contractions run
    specs = [CircuitSpec(9, 4, 7, 1), CircuitSpec(14, 8, 12, 2), CircuitSpec(30, 16, 9, 3),
             CircuitSpec(21, 32, 14, 4), CircuitSpec(12, 64, 16, 5), CircuitSpec(2, 2, 3, 6),
             CircuitSpec(14, 100, 14, 7), CircuitSpec(16, 128, 16, 8)]
    cut = 1e-10
    run = MultiRun(specs, cutoff=cut, precision=precision, max_tasks=40)
    run.run()
    run.check()
    assert run.chains[6].max_bond() == 100 and run.chains[7].max_bond() == 128
    logs = run.norm_logs()
    for c, s in enumerate(specs):
        ref = oracle_run(s, cut)
        host = run.chains[c].to_host()
        got = mps_ref.MPS(s.n, 1, 0)
        got.sites, got.norm_log = host.sites, logs[c]
        if precision == 'fp64':
            assert host.bond_dims() == ref.bond_dims(), c
            assert abs(logs[c] - ref.norm_log) < 1e-10 * max(1, abs(ref.norm_log))
        ov = mps_ref.overlap(got, ref)
        fid = abs(ov) ** 2 / (mps_ref.overlap(got, got).real * mps_ref.overlap(ref, ref).real)
        assert 1 - fid < (1e-10 if precision == 'fp64' else 1e-6), (c, 1 - fid)


def test_graph_replay_matches_eager_run():
    """The whole multi-circuit run recorded as graphs (one per round of updates): every
    replay (from the same start, restored by eager copies between replays) equals the
    eager run bit for bit."""
    import torch
    from qsim_amd.mps_multi import CircuitSpec, MultiRun
    specs = [CircuitSpec(9, 4, 7, 1), CircuitSpec(14, 8, 12, 2), CircuitSpec(30, 16, 9, 3),
             CircuitSpec(21, 32, 14, 4), CircuitSpec(12, 64, 16, 5), CircuitSpec(16, 100, 14, 6),
             CircuitSpec(18, 128, 10, 8)]
    eager = MultiRun(specs, cutoff=1e-10, precision='fp32', max_tasks=25)
    eager.run()
    eager.check()
    gr = MultiRun(specs, cutoff=1e-10, precision='fp32', max_tasks=25)
    start = [(ch.sites.clone(), ch.bond.clone()) for ch in gr.chains]
    gr.build_graph()
    assert len(gr.g_chunks) > max(s.depth for s in specs)      # several chunks per round
    for rep in range(3):
        if rep:
            for ch, (s0, b0) in zip(gr.chains, start):
                ch.sites.copy_(s0)
                ch.bond.copy_(b0)
            gr.g_status.zero_()
            torch.cuda.synchronize()
        gr.replay()
        torch.cuda.synchronize()
        gr.check_graph()
        for c in range(len(specs)):
            assert torch.equal(gr.chains[c].bond, eager.chains[c].bond), (rep, c)
            assert torch.equal(gr.chains[c].sites, eager.chains[c].sites), (rep, c)
        assert np.array_equal(gr.norm_logs(), eager.norm_logs()), rep
