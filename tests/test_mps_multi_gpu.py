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
    specs = [CircuitSpec(9, 4, 7, 1), CircuitSpec(14, 8, 12, 2), CircuitSpec(30, 16, 9, 3),
             CircuitSpec(21, 32, 14, 4), CircuitSpec(12, 64, 16, 5), CircuitSpec(2, 2, 3, 6)]
    cut = 1e-10
    run = MultiRun(specs, cutoff=cut, precision=precision, max_tasks=40)
    run.run()
    run.check()
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
