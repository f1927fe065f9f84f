"""qs_mps_update_batch_static (the allocation-free two-site batch, static device
scratch) against qs_mps_update_batch with a caller workspace: bit-identical sites,
bond dimensions, norm_log increments and statuses, on consecutive calls whose batches
shrink, for every SVD size class, FP64 and FP32-MFMA; batches beyond the scratch size
are refused with an error and change nothing."""

import numpy as np
import pytest

from qsim_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _pair(n, chi, cutoff, precision):
    from qsim_amd.mps import MPSChain
    return (MPSChain(n, chi, cutoff, precision=precision),
            MPSChain(n, chi, cutoff, precision=precision))


def _same(a, b):
    import torch
    torch.cuda.synchronize()
    assert torch.equal(a.bond, b.bond)
    assert torch.equal(a.sites, b.sites)
    assert len(a._dlogs) == len(b._dlogs)
    for x, y in zip(a._dlogs, b._dlogs):
        assert torch.equal(x, y)
    assert np.array_equal(a.last_sweeps(), b.last_sweeps())


@pytest.mark.parametrize('chi,precision', [(16, 'fp64'), (32, 'fp32'), (64, 'fp32'),
                                           (64, 'fp64'), (100, 'fp64')])
def test_static_scratch_batches_equal_workspace_batches(chi, precision):
    n, depth = 24, 10
    ref, got = _pair(n, chi, 1e-12, precision)
    gates = workloads.brickwall_gates(n, depth, seed=chi, kind='haar')
    layers = {}
    for lay, i, g in gates:
        layers.setdefault(lay, []).append((i, g))
    for lay in sorted(layers):
        batch = layers[lay]
        # two calls in a row, the second smaller: the whole layer, then what is left
        # after its first half
        for part in (batch[:max(1, len(batch) // 2 + 2)], batch[len(batch) // 2 + 2:]):
            if not part:
                continue
            ref.apply_2q_batch(part)
            got.apply_2q_batch(part, allocation_free=True)
            _same(ref, got)
    ref.check_status()
    got.check_status()
    assert max(got.bond_dims()) > min(chi, 8)


def test_static_scratch_largest_batch_then_smaller_and_refusal():
    import torch
    chi = 64
    lib = _lib.lib()
    cap = _lib.QS_MPS_STATIC_SCRATCH_BYTES
    kmax = max(k for k in range(1, 400) if lib.qs_mps_workspace_bytes(k, chi) <= cap)
    assert kmax == 127
    n = 2 * kmax + 3
    ref, got = _pair(n, chi, 1e-10, 'fp32')
    rng = np.random.default_rng(5)
    big = [(2 * j, workloads.near_identity_4x4(rng)) for j in range(kmax)]
    small = [(2 * j + 1, workloads.near_identity_4x4(rng)) for j in range(9)]
    ref.apply_2q_batch(big)
    got.apply_2q_batch(big, allocation_free=True)
    _same(ref, got)
    ref.apply_2q_batch(small)
    got.apply_2q_batch(small, allocation_free=True)
    _same(ref, got)
    # one task more than the scratch holds: an error, and the chain is left as it was
    snap_sites, snap_bond = got.sites.clone(), got.bond.clone()
    over = [(2 * j, workloads.near_identity_4x4(rng)) for j in range(kmax + 1)]
    with pytest.raises(_lib.QsError) as err:
        got.apply_2q_batch(over, allocation_free=True)
    assert err.value.status == _lib.QS_E_RANGE
    assert 'static scratch' in str(err.value)
    torch.cuda.synchronize()
    assert torch.equal(got.sites, snap_sites) and torch.equal(got.bond, snap_bond)
