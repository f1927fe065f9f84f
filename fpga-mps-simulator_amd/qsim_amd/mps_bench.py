"""MPS leg of bench.py: two-site updates/s at chi=64.

Workload: 1000 qubits, |0..0>, brick-wall depth 40 of exp(-0.1 i H) gates (H seeded
random Hermitian), chi_max 64, cutoff 1e-10: 20 x 500 + 20 x 499 = 19 980 updates.

  fp32_batched : every layer is one batch (complex FP32 contraction on MFMA, FP64 SVD)
  fp64_batched : the same batches with the FP64 contraction
  fp64_plain   : one update per launch sequence, FP64 (mps_core semantics, gate by gate),
                 timed on a bounded sample: the updates of layers `plain_layers` after
                 replaying the earlier layers untimed, so the sample sees the same bond
                 dimensions as the batched runs at those layers

Per-layer times come from HIP events on the stream the kernels run on.  Before timing,
the SVD team width of the chi-64 class is tuned (mps_tune) and reported.

run_multi(): many independent circuits (n ~ U{50..2000}, chi_max in {16, 32, 64, 128},
depth ~ U{10..60}, exp(-0.1iH) gates, complex-FP32 contraction + FP64 SVD), all rounds
recorded as one graph and replayed; updates/s over the replay.
"""

from __future__ import annotations

import math
import time
from typing import Dict, List

import numpy as np

N_SITES = 1000
DEPTH = 40
CHI = 64
CUTOFF = 1e-10
SEED = 42
SHOTS = 1024
MFMA_FP32_PEAK_TFLOPS = 157.3        # MI355X dense FP32 matrix
FP64_PEAK_TFLOPS = 78.6              # MI355X FP64 (vector and matrix)


def contraction_flops(cl: int, cm: int, cr: int) -> float:
    """Algorithmic real flops of one update's contraction: 32 chi_l chi_r (chi_m + 4)."""
    return 32.0 * cl * cr * (cm + 4)


def svd_qr_flops(nrow: int, ncol: int) -> float:
    """Algorithmic real flops of the preconditioning Householder QR of the nrow x ncol
    complex matrix theta^H (nrow = 2 chi_r, ncol = 2 chi_l): for each of the
    k = min(nrow, ncol) reflections, every later column forms v^H x and x - t v over the
    rows at and below the diagonal (16 real flops per row and column): about
    8 (2 nrow ncol k - (nrow + ncol) k^2 + 2 k^3 / 3)."""
    k = min(nrow, ncol)
    return 8.0 * (2.0 * nrow * ncol * k - (nrow + ncol) * k * k + 2.0 * k ** 3 / 3.0)


def svd_update_flops(cl: int, cr: int, sweeps: int) -> float:
    """Algorithmic real flops of one update's SVD: the QR of theta^H (2 chi_r x 2 chi_l)
    plus `sweeps` one-sided Jacobi sweeps over L = R^H (2 chi_l rows, min(2 chi_l, 2 chi_r)
    columns)."""
    m, n = 2 * cl, 2 * cr
    return svd_qr_flops(n, m) + sweeps * svd_sweep_flops(m, min(m, n))


def svd_sweep_flops(m: int, ncol: int) -> float:
    """Algorithmic real flops of one one-sided Jacobi sweep over an m x ncol complex
    matrix: every column pair forms a^H b (8 m) and is rotated (two columns, 10 m each:
    real cosine times a column, complex sine times the other, and their sum), and every
    column's squared norm is formed once (4 m)."""
    pairs = ncol * (ncol - 1) / 2
    return pairs * 28.0 * m + 4.0 * m * ncol


def _layers(gates):
    out: Dict[int, list] = {}
    for lay, i, g in gates:
        out.setdefault(lay, []).append((i, g))
    return [out[k] for k in sorted(out)]


def run_batched(layers, precision: str):
    import torch
    from .mps import MPSChain
    ch = MPSChain(N_SITES, CHI, CUTOFF, precision=precision)
    stream = torch.cuda.current_stream()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(layers) + 1)]
    flops = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k, batch in enumerate(layers):
        ch.apply_2q_batch(batch)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_layer = [evs[k].elapsed_time(evs[k + 1]) / 1e3 for k in range(len(layers))]
    ch.check_status()
    return ch, per_layer, wall


def run_contraction_timed(layers):
    """The FP32 batched run again with every contraction and SVD launch timed by HIP
    events, and the bond dimensions read before each layer, so the contraction's
    algorithmic flops (32 chi_l chi_r (chi_m + 4) per update) are exact."""
    import ctypes
    import torch
    from . import _lib
    from .mps import MPSChain
    ch = MPSChain(N_SITES, CHI, CUTOFF, precision='fp32')
    flops = 0.0
    svd_flops, sweeps = 0.0, []
    _lib.call('qs_mps_time_kernels', 1)
    try:
        for batch in layers:
            b = ch.bond.cpu().numpy().astype(np.float64)
            ch.apply_2q_batch(batch)
            sw = ch.last_sweeps()
            for (i, _), nsw in zip(batch, sw):
                flops += contraction_flops(b[i], b[i + 1], b[i + 2])
                svd_flops += svd_update_flops(int(b[i]), int(b[i + 2]), int(nsw))
            sweeps.extend(int(v) for v in sw)
        torch.cuda.synchronize()
        cap = 8 * len(layers)
        kind = (ctypes.c_int32 * cap)()
        tasks = (ctypes.c_int32 * cap)()
        ms = (ctypes.c_double * cap)()
        cnt = ctypes.c_int(0)
        _lib.call('qs_mps_kernel_times', kind, tasks, ms, cap, ctypes.byref(cnt))
    finally:
        _lib.call('qs_mps_time_kernels', 0)
    ch.check_status()
    n = min(cnt.value, cap)
    t_con = sum(ms[k] for k in range(n) if kind[k] == 0) / 1e3
    t_svd = sum(ms[k] for k in range(n) if kind[k] == 2) / 1e3
    return flops, t_con, t_svd, svd_flops, sweeps


def run_plain_sample(layers, sample_layers: List[int]):
    """FP64, one update per launch; earlier layers replayed batched (untimed)."""
    import torch
    from .mps import MPSChain
    ch = MPSChain(N_SITES, CHI, CUTOFF, precision='fp64')
    first = min(sample_layers)
    for batch in layers[:first]:
        ch.apply_2q_batch(batch)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = 0
    e0.record(stream)
    for k in sample_layers:
        for upd in layers[k]:
            ch.apply_2q_batch([upd])
            n += 1
    e1.record(stream)
    torch.cuda.synchronize()
    ch.check_status()
    return n, e0.elapsed_time(e1) / 1e3


def run(plain_layers=(30, 31)) -> dict:
    import torch
    from . import workloads
    gates = workloads.brickwall_gates(N_SITES, DEPTH, SEED)
    layers = _layers(gates)
    nupd = sum(len(b) for b in layers)
    # warm-up (kernel load, allocator) on a short chain
    from .mps import MPSChain
    w = MPSChain(8, CHI, CUTOFF, precision='fp32')
    w.apply_2q_batch([(0, gates[0][2]), (2, gates[1][2])])
    torch.cuda.synchronize()

    from . import mps_tune
    tuned = mps_tune.autotune(classes=(CHI,), precision='fp32')
    ch32, t32, wall32 = run_batched(layers, 'fp32')
    ch64, t64, wall64 = run_batched(layers, 'fp64')
    n_plain, t_plain = run_plain_sample(layers, list(plain_layers))
    flops, t_con, t_svd, svd_flops, sweeps = run_contraction_timed(layers)
    # shot sampling of the final 1000-qubit state (right environments + 1024 shots)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ch64.sample(SHOTS, None, np.random.default_rng(SEED))
    torch.cuda.synchronize()
    t_samp = time.perf_counter() - t0
    bonds = ch64.bond_dims()
    # FP32-contraction state against the FP64 one, by transfer matrices on the device
    fid = ch32.fidelity(ch64)
    dz = float(np.max(np.abs(ch32.z_expectations() - ch64.z_expectations())))
    s32 = sum(t32)
    s64 = sum(t64)
    samp32 = sum(t32[k] for k in plain_layers)
    samp64 = sum(t64[k] for k in plain_layers)
    n_samp = sum(len(layers[k]) for k in plain_layers)
    rate32 = nupd / s32
    rate_plain = n_plain / t_plain
    return {
        'workload': f'{N_SITES} qubits, brick-wall depth {DEPTH} of exp(-0.1iH), chi_max {CHI}, '
                    f'cutoff {CUTOFF:g}: {nupd} two-site updates',
        'updates': nupd,
        'fp32_mfma_batched_updates_per_s': rate32,
        'fp64_batched_updates_per_s': nupd / s64,
        'fp64_plain_updates_per_s': rate_plain,
        'fp64_plain_sample': f'{n_plain} updates of layers {list(plain_layers)}, one per launch',
        'fp32_batched_updates_per_s_same_layers': n_samp / samp32,
        'fp64_batched_updates_per_s_same_layers': n_samp / samp64,
        'speedup_fp32_batched_vs_fp64_plain_same_layers': (n_samp / samp32) / rate_plain,
        'fp32_seconds': s32, 'fp64_seconds': s64,
        'max_bond': max(bonds), 'mean_bond': float(np.mean(bonds)),
        'norm_log_fp64': ch64.norm_log, 'norm_log_fp32': ch32.norm_log,
        'infidelity_fp32_vs_fp64': float(1 - fid),
        'max_abs_z_error_fp32_vs_fp64': dz,
        'contraction': {
            'kernel': 'mps_contract_f32_mfma_kernel',
            'algorithmic_tflop': flops / 1e12,
            'seconds': t_con,
            'tflops': flops / t_con / 1e12,
            'mfma_peak_tflops': MFMA_FP32_PEAK_TFLOPS,
            'mfma_frac': flops / t_con / 1e12 / MFMA_FP32_PEAK_TFLOPS,
            'svd_seconds_same_run': t_svd,
        },
        'svd': {
            'kernel': 'mps_svd_kernel<64,*>',
            'model': 'per update: Householder QR of theta^H (2 chi_r x 2 chi_l) + sweeps x '
                     '(ncol(ncol-1)/2 x 28 m + 4 m ncol) real flops of one-sided Jacobi on '
                     'L = R^H, m = 2 chi_l rows, ncol = min(2 chi_l, 2 chi_r) columns',
            'mean_sweeps_per_update': float(np.mean(sweeps)),
            'algorithmic_tflop': svd_flops / 1e12,
            'seconds': t_svd,
            'tflops': svd_flops / t_svd / 1e12,
            'fp64_peak_tflops': FP64_PEAK_TFLOPS,
            'fp64_frac': svd_flops / t_svd / 1e12 / FP64_PEAK_TFLOPS,
            'share_of_update_kernel_time': t_svd / (t_svd + t_con),
        },
        'svd_team_width': {str(c): r['lanes'] for c, r in tuned.items()},
        'sampling_shots_per_s': SHOTS / t_samp,
        'sampling_sample': f'{SHOTS} shots of the final {N_SITES}-qubit chi {CHI} state, '
                           f'right environments included, {t_samp:.3f} s',
    }


def run_multi(ncirc: int = 256, seed: int = 7, precision: str = 'fp32',
              max_tasks: int = 4096, quiet: bool = False) -> dict:
    """The batched multi-circuit workload (the first `ncirc` circuits of the seeded
    256-circuit list), planned and recorded as graphs, one per round of updates."""
    import torch
    from . import mps_tune
    from .mps_multi import MultiRun, random_specs
    specs = random_specs(256, seed=seed, chis=(16, 32, 64, 128))[:ncirc]
    tuned = mps_tune.autotune(precision=precision)
    import sys
    run = MultiRun(specs, cutoff=1e-10, precision=precision, max_tasks=max_tasks)
    t0 = time.perf_counter()

    def planned(r, nr):
        if not quiet and (r % 10 == 9 or r == nr - 1):
            print(f'multi: planned round {r + 1}/{nr}, {time.perf_counter() - t0:.0f} s',
                  file=sys.stderr, flush=True)
    run.build_graph(progress=planned)
    plan_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    def replayed(k, nk):
        # one synchronisation per ten rounds, outside the replayed graphs: a progress line
        if not quiet and k % 10 == 9:
            torch.cuda.synchronize()
            print(f'multi: replayed round {k + 1}/{nk}, {time.perf_counter() - t1:.0f} s',
                  file=sys.stderr, flush=True)
    e0.record(stream)
    run.replay(progress=replayed)
    e1.record(stream)
    torch.cuda.synchronize()
    run.check_graph()
    secs = e0.elapsed_time(e1) / 1e3
    per_class = {}
    upd = np.bincount(run.owner[:run._next], minlength=len(specs))
    for c in (16, 32, 64, 128):
        idx = [i for i, s in enumerate(specs) if s.chi_max == c]
        per_class[str(c)] = {'circuits': len(idx), 'updates': int(sum(upd[i] for i in idx)),
                             'max_bond': max((run.chains[i].max_bond() for i in idx), default=0)}
    return {
        'workload': f'{ncirc} independent circuits, n ~ U{{50..2000}}, chi_max in '
                    f'{{16, 32, 64, 128}}, depth ~ U{{10..60}}, exp(-0.1iH) gates, cutoff 1e-10, '
                    f'{precision} contraction + FP64 SVD',
        'updates': run._next,
        'updates_per_s': run._next / secs,
        'seconds': secs,
        'graph_chunks': len(run.g_chunks),
        'graphs': len(run.graphs),
        'plan_and_record_seconds': plan_s,
        'classes': per_class,
        'svd_team_width': {str(c): r['lanes'] for c, r in tuned.items()},
    }
