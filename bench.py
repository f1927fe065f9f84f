#!/usr/bin/env python3
"""Benchmark of the MI355X gate-application paths; prints ONE JSON line.

Headline (`value`): 28-qubit complex128 statevector sweep, gates/s, on the
workload "20 layers of U3 on every qubit + CZ ladder, then QFT-28" (1520
logical gates) with the state resident in HBM.  A step applies the whole
circuit once: one launch of a plan built before the timed region
(qs_sv_plan_create; its host time is reported as `config.plan_ms`), whose passes run
kernels compiled for the plan (qs_sv_plan_specialize, hipRTC; the compile time is
reported as `config.jit_ms`; `--interpreter` runs the op-stream interpreter instead).
The line also carries:
  roofline     : algorithmic HBM bytes of the dominant kernel (the fused pass:
                 read + write of the 4 GiB state) per launch over its average
                 duration, timed with HIP events around every pass launch of the
                 timed steps, against 8 TB/s; `traffic` is the measured HBM bytes
                 per launch from the committed rocprofv3 counter summary,
  cpu_baseline : the NumPy oracle's pair updates on the host, on a bounded
                 sample of the same 28-qubit workload (and, under `mps`, the MPS
                 oracle's two-site updates at chi=64 on a sample of one layer),
  mps          : two-site MPS updates/s at chi=64 (1000 qubits, brick-wall of
                 exp(-0.1iH) gates) for the batched FP32-MFMA path and the
                 plain FP64 path (one update per launch), when built.

Also measured live, as a reference for the roofline: `inplace_rmw_gbps`, an in-place
read-modify-write of the same 4 GiB state by torch's elementwise kernel (x *= 1),
i.e. what a plain streaming kernel reaches on this HBM.

`--multi NCIRC` prints a separate line for the batched multi-circuit MPS workload
(NCIRC random circuits, chi_max 16..128, all rounds recorded as one graph).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mps]
                       [--multi NCIRC]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'fpga-mps-simulator_amd')
sys.path.insert(0, PKG)

METRIC = "MPS 2q updates/s @chi=64 (FP32-MFMA vs FP64); 28q SV gates/s + HBM GB/s"
HBM_PEAK_GBPS = 8000.0
SV_QUBITS = 28
SV_DEPTH = 20
SV_SEED = 1234


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-mps', action='store_true', help='skip the MPS leg')
    ap.add_argument('--cpu-gates', type=int, default=6,
                    help='logical gates of the workload timed on the CPU')
    ap.add_argument('--multi', type=int, default=0, metavar='NCIRC',
                    help='instead of the headline line, print one JSON line for NCIRC '
                         'independent MPS circuits (chi_max 16..128) replayed as one graph')
    ap.add_argument('--multi-sample', type=int, default=24, metavar='NCIRC',
                    help='circuits of the many-circuit MPS workload run inside the default line')
    ap.add_argument('--pass-table', metavar='PATH', default=None,
                    help='also write a per-pass table (mean kernel time over the timed steps, '
                         'tile qubits, stages and op counts by kind) as markdown to PATH')
    ap.add_argument('--calibration-copy', action='store_true',
                    help='after the SV leg, copy the 4 GiB state once (a copy kernel of '
                         'known size for scaling the HBM byte counters of a profile)')
    ap.add_argument('--sv-sizes', default='', metavar='N,N,..',
                    help='instead of the headline line, print one JSON line per state size: '
                         'the statevector workload structure (20 x [U3 all + CZ ladder] + '
                         'QFT) on N qubits, per-plan kernels, --steps timed launches')
    ap.add_argument('--interpreter', action='store_true',
                    help='run the plan on the op-stream interpreter kernel instead of '
                         'kernels compiled for the plan')
    return ap.parse_args()


def sv_size(nq: int, steps: int, warmup: int) -> dict:
    """The statevector workload's structure at nq qubits: plan, per-plan kernels, `steps`
    timed launches of the prebuilt plan (HIP events on the launch stream)."""
    import torch
    from qsim_amd import workloads
    from qsim_amd.sv import StateVector, gates_to_records

    gates = workloads.sv_benchmark_circuit(nq, SV_DEPTH, SV_SEED)
    recs, _ = gates_to_records(gates, nq)
    sv = StateVector(nq, 0)
    t0 = time.perf_counter()
    plan = sv.plan(recs)
    plan_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan.specialize()
    jit_s = time.perf_counter() - t0
    for _ in range(warmup):
        sv.launch(plan)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        sv.launch(plan)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    per_pass_s = ms / 1e3 / plan.npasses
    out = {'workload': f'{nq}q statevector: {SV_DEPTH}x[U3 all qubits + CZ ladder] + '
                       f'QFT-{nq} ({len(gates)} gates), complex128, state resident in HBM',
           'qubits': nq, 'gates': len(gates), 'passes_per_step': plan.npasses,
           'steps': steps, 'ms_per_step': ms, 'gates_per_s': len(gates) / (ms / 1e3),
           'per_pass_ms': per_pass_s * 1e3,
           'per_pass_gbps': 2 * 16 * (1 << nq) / per_pass_s / 1e9,
           'plan_ms': plan_s * 1e3, 'jit_ms': jit_s * 1e3, 'norm_sq_after': sv.norm_sq()}
    del sv, plan
    torch.cuda.empty_cache()
    return out


def sv_leg(steps: int, warmup: int, calibration_copy: bool = False,
           specialize: bool = True):
    import torch
    from qsim_amd import workloads
    from qsim_amd.sv import StateVector, gates_to_records, pass_times, time_passes

    gates = workloads.sv_benchmark_circuit(SV_QUBITS, SV_DEPTH, SV_SEED)
    recs, _ = gates_to_records(gates, SV_QUBITS)
    sv = StateVector(SV_QUBITS, 0)
    stream = torch.cuda.current_stream()
    # plan once (host planner + upload), outside the timed region and reported apart
    t0 = time.perf_counter()
    plan = sv.plan(recs)
    plan_s = time.perf_counter() - t0
    jit_s, jit_kernels = 0.0, 0
    if specialize:
        t0 = time.perf_counter()
        jit_kernels = plan.specialize()
        jit_s = time.perf_counter() - t0
    npass = plan.npasses
    for _ in range(warmup):
        sv.launch(plan)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the library brackets every pass-kernel launch with HIP events on this stream
    time_passes(True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        sv.launch(plan)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = pass_times()
    time_passes(False)
    gpu_ms = ev0.elapsed_time(ev1)
    norm = sv.norm_sq()
    # reference: a plain in-place streaming kernel over the same buffer
    r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sv.tensor.mul_(1.0)
    r0.record(stream)
    for _ in range(3):
        sv.tensor.mul_(1.0)
    r1.record(stream)
    torch.cuda.synchronize()
    rmw_gbps = 3 * 2 * 16 * (1 << SV_QUBITS) / (r0.elapsed_time(r1) / 1e3) / 1e9
    if calibration_copy:
        sv.tensor.clone()
        torch.cuda.synchronize()
    bytes_per_pass = 2 * 16 * (1 << SV_QUBITS)
    if len(kernel_ms) != steps * npass:
        raise RuntimeError(f'expected {steps * npass} pass timings, got {len(kernel_ms)}')
    per_launch_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    per_pass_ms = [sum(kernel_ms[k * npass + p] for k in range(steps)) / steps
                   for p in range(npass)]
    return {
        'per_pass_ms': per_pass_ms, 'records': recs, 'plan_s': plan_s,
        'jit_s': jit_s, 'jit_kernels': jit_kernels, 'specialized': specialize,
        'plan_device_bytes': plan.device_bytes,
        'wall_s': wall, 'gpu_ms': gpu_ms, 'npass': npass, 'ngates': len(gates),
        'per_launch_s': per_launch_s, 'bytes_per_pass': bytes_per_pass,
        'norm_sq': norm, 'inplace_rmw_gbps': rmw_gbps,
    }


_OP_NAMES = {0: 'U1', 1: 'U2', 2: 'DIAG1', 3: 'DIAG2', 4: 'U1X', 5: 'PHASE', 6: 'LAYOUT',
             7: 'SIGN', 8: 'U1S', 9: 'SWAP', 10: 'SCALE', 11: 'SIGN_REG'}


def write_pass_table(path: str, sv: dict, cmd: str) -> None:
    """Markdown table of the fused passes of one step: mean kernel time over the timed
    steps, bandwidth, tile qubits, stages and ops by kind (from the plan JSON)."""
    from qsim_amd.sv import plan_json
    plan = plan_json(SV_QUBITS, sv['records'])['passes']
    bpp = sv['bytes_per_pass']
    lines = [f'# Fused passes of the 28-qubit statevector step ({len(plan)} passes)', '',
             f'command: `{cmd}`', '',
             'Per pass: mean duration of the pass kernel over the timed steps (HIP events on '
             'the launch stream), algorithmic bandwidth (2 x 16 x 2^28 B per pass), tile '
             'qubits, stages and op counts by kind (ROT/ROTS = rotations, SIGN = +-1 phase '
             'groups, PHASE = complex phase groups).', '',
             '| pass | ms | TB/s | frac of 8 TB/s | tile qubits | stages | ops |',
             '|---:|---:|---:|---:|---|---:|---|']
    total = 0.0
    for i, (ms, P) in enumerate(zip(sv['per_pass_ms'], plan)):
        total += ms
        kinds = {}
        for S in P['stages']:
            for G in S['gates']:
                k = G['kind']
                nm = 'ROTS' if k >= 32 else 'ROT' if k >= 16 else _OP_NAMES.get(k, str(k))
                kinds[nm] = kinds.get(nm, 0) + 1
        ops = ' '.join(f'{k}:{v}' for k, v in sorted(kinds.items()))
        tq = ','.join(str(q) for q in P['tq'])
        tbs = bpp / (ms / 1e3) / 1e12
        lines.append(f'| {i} | {ms:.3f} | {tbs:.2f} | {tbs / 8.0:.3f} | {tq} | '
                     f'{len(P["stages"])} | {ops} |')
    lines += ['', f'sum of pass means: {total:.2f} ms per step; mean per pass '
              f'{total / len(plan):.3f} ms = {bpp / (total / len(plan) / 1e3) / 1e12:.2f} TB/s']
    with open(path, 'w') as f:
        f.write('\n'.join(lines) + '\n')


def cpu_model() -> str:
    """The host CPU's model name (first `model name` line of /proc/cpuinfo)."""
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_omp_leg(nlayer_gates: int = 60):
    """The oracle's C/OpenMP pair updates (oracle/sv_kernel_omp.c) on a bounded sample of
    the 28-qubit workload: the first `nlayer_gates` gates of the layered part (U3s and the
    CZ ladder), plus four QFT controlled phases, lowered as the reference lowers them."""
    import numpy as np
    from oracle import sv_omp, sv_ref
    from qsim_amd import workloads
    gates = workloads.sv_benchmark_circuit(SV_QUBITS, SV_DEPTH, SV_SEED)
    sample = gates[:nlayer_gates] + [g for g in gates if g[0] == 'cp'][:4]
    ops = sv_ref.lower(sample)
    re = np.zeros(1 << SV_QUBITS)
    im = np.zeros(1 << SV_QUBITS)
    re[0] = 1.0
    t0 = time.perf_counter()
    sv_omp.apply_kernel_ops(re, im, SV_QUBITS, ops)
    dt = time.perf_counter() - t0
    nthr = sv_omp.threads()
    return {'value': len(sample) / dt, 'unit': 'gates/s', 'cores': nthr, 'kind': 'port',
            'sample': f'{len(sample)} logical gates ({len(ops)} kernel ops after the reference '
                      f'lowering) of the 28q workload, C/OpenMP pair updates on split re/im '
                      f'arrays (oracle/sv_kernel_omp.c), {nthr} threads, {dt:.1f} s'}


def cpu_leg(ngates: int):
    """NumPy pair updates of the oracle on the first `ngates` logical gates at 28 qubits."""
    import numpy as np
    from oracle import sv_ref
    from qsim_amd import workloads
    gates = workloads.sv_benchmark_circuit(SV_QUBITS, SV_DEPTH, SV_SEED)
    # a bounded sample: U3s of the first layer and two CZs, plus one cp from the QFT
    sample = gates[:ngates - 3] + [g for g in gates if g[0] == 'cz'][:2] + \
        [g for g in gates if g[0] == 'cp'][:1]
    state = np.zeros(1 << SV_QUBITS, dtype=np.complex128)
    state[0] = 1.0
    ops = sv_ref.lower(sample)
    t0 = time.perf_counter()
    sv_ref.apply_kernel_ops(state, SV_QUBITS, ops)
    dt = time.perf_counter() - t0
    return {'value': len(sample) / dt, 'unit': 'gates/s', 'cores': 1, 'kind': 'port',
            'sample': f'{len(sample)} logical gates ({len(ops)} kernel ops after the '
                      f'reference lowering) of the 28q workload, NumPy pair updates, '
                      f'{dt:.1f} s'}


def cpu_mps_leg(nsample: int = 500, layer: int = 30):
    """The MPS oracle (NumPy/SciPy restatement of mps_core) on the host: `nsample`
    two-site updates of layer `layer` of the chi=64 workload, starting from the state the
    GPU's FP64 path reaches after the earlier layers (copied to the host)."""
    import numpy as np
    import torch
    from oracle import mps_ref
    from qsim_amd import mps_bench, workloads
    from qsim_amd.mps import MPSChain
    gates = workloads.brickwall_gates(mps_bench.N_SITES, mps_bench.DEPTH, mps_bench.SEED)
    layers = mps_bench._layers(gates)
    ch = MPSChain(mps_bench.N_SITES, mps_bench.CHI, mps_bench.CUTOFF, precision='fp64')
    for batch in layers[:layer]:
        ch.apply_2q_batch(batch)
    torch.cuda.synchronize()
    host = ch.to_host()
    ref = mps_ref.MPS(mps_bench.N_SITES, mps_bench.CHI, mps_bench.CUTOFF)
    ref.sites = host.sites
    sample = layers[layer][:nsample]
    import multiprocessing as mp
    from threadpoolctl import threadpool_limits
    # 128 x 128 SVDs run fastest with one BLAS thread each; the updates of one layer are
    # independent (disjoint bonds), so the parallel row hands one update at a time to a
    # pool of worker processes (one BLAS thread each), the two sites travelling with it
    nproc = min(16, os.cpu_count() or 1)
    items = [(ref.sites[i], ref.sites[i + 1], g, mps_bench.CHI, mps_bench.CUTOFF)
             for i, g in sample]
    saved = {k: os.environ.get(k) for k in ('OPENBLAS_NUM_THREADS', 'OMP_NUM_THREADS')}
    os.environ.update({'OPENBLAS_NUM_THREADS': '1', 'OMP_NUM_THREADS': '1'})
    try:
        with mp.get_context('spawn').Pool(nproc) as pool:
            pool.map(mps_ref.update_pair, items[:nproc], chunksize=1)     # warm-up
            t0 = time.perf_counter()
            pool.map(mps_ref.update_pair, items, chunksize=1)
            dtp = time.perf_counter() - t0
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for i, g in sample:
            ref.apply_adjacent_2q(i, g)
        dt = time.perf_counter() - t0
    nthr = nproc
    # host sampling of the same chain: the right environments of the last `nenv` sites
    # (the reference's einsum, one per site, the same chi 64 x chi 64 work at every bulk
    # site; the few cheaper edge sites are all inside the window), scaled to the chain,
    # plus `shots` shots over environments of the same shapes
    nenv, shots = 200, 16
    t1 = time.perf_counter()
    envs = ref.right_environments(stop=ref.n - nenv)
    te = (time.perf_counter() - t1) * ref.n / nenv
    flat = [np.eye(a.shape[0], dtype=np.complex128) for a in ref.sites] + [envs[ref.n]]
    t1 = time.perf_counter()
    ref.sample(shots, None, np.random.default_rng(mps_bench.SEED), envs=flat)
    tshot = time.perf_counter() - t1
    ts = te + tshot
    return {'value': len(sample) / dtp, 'unit': 'updates/s',
            'cores': nthr, 'kind': 'port',
            'sample': f'{len(sample)} two-site updates (chi 64, layer {layer}) of the '
                      f'1000-qubit workload, NumPy/SciPy (LAPACK gesdd), one update per '
                      f'task over {nthr} worker processes, one BLAS thread each, {dtp:.2f} s',
            'single_thread': {'value': len(sample) / dt, 'unit': 'updates/s', 'cores': 1,
                              'sample': f'the same updates in order on one thread, {dt:.1f} s'},
            'sampling_shots_per_s': shots / ts,
            'sampling_sample': f'{shots} shots of the 1000-qubit chi 64 chain plus its right '
                               f'environments (the reference einsum formulation), '
                               f'{ts:.1f} s: environments of the last {nenv} sites timed '
                               f'and scaled to {ref.n} ({te:.1f} s), the shots over '
                               f'same-shape environments ({tshot:.1f} s)'}


def gradients_leg():
    """C5 gradients on the GPU: seconds per full gradient of the TFIM energy (J = h = 1,
    open chain) of the hardware-efficient ansatz (4 x [Ry all, Rz all, CZ ladder]), params
    ~ U[-pi, pi]: the statevector adjoint at 24 qubits (192 parameters) and the MPS reverse
    sweep at 200 qubits, chi_max 32 (1600 parameters)."""
    import numpy as np
    import torch
    from qsim_amd.adjoint import TFIMEnergy, hardware_efficient_ansatz
    from qsim_amd.mps_grad import MPSEnergyGrad
    rng = np.random.default_rng(2024)
    out = {}
    c = hardware_efficient_ansatz(24, 4)
    th = rng.uniform(-math.pi, math.pi, c.nparams)
    mod = TFIMEnergy(c)
    mod.energy_and_grad(th)                       # warm-up (allocation, kernel load)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        e, g = mod.energy_and_grad(th)
    torch.cuda.synchronize()
    out['sv_24q'] = {'params': c.nparams, 'seconds_per_gradient': (time.perf_counter() - t0) / reps,
                     'energy': e, 'method': 'adjoint reverse sweep (one forward, two reverse '
                                             'passes, one Pauli reduction per rotation block)'}
    del mod
    c2 = hardware_efficient_ansatz(200, 4)
    th2 = rng.uniform(-math.pi, math.pi, c2.nparams)
    eng = MPSEnergyGrad(c2, chi_max=32, cutoff=1e-12)
    eng.energy_and_grad(th2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e2, g2 = eng.energy_and_grad(th2)
    torch.cuda.synchronize()
    out['mps_200q_chi32'] = {'params': c2.nparams,
                             'seconds_per_gradient': time.perf_counter() - t0, 'energy': e2,
                             'method': 'GPU forward with tape + reverse sweep through the '
                                       'contractions and truncated SVDs'}
    return out


def cpu_gradients_leg():
    """The oracle's finite-difference gradient cost for the same C5 circuits, extrapolated
    from a bounded sample: central differences need 2 P energy evaluations; one energy is
    timed on a sample of the circuit (the first gates) and scaled to the whole circuit."""
    import numpy as np
    from oracle import mps_ref, sv_ref
    rng = np.random.default_rng(2024)
    out = {}
    n, layers = 24, 4
    th = rng.uniform(-math.pi, math.pi, 2 * n * layers)
    gates = sv_ref.hea_gates(n, th, layers)
    ns = 30
    state = np.zeros(1 << n, dtype=np.complex128)
    state[0] = 1.0
    t0 = time.perf_counter()
    sv_ref.apply_kernel_ops(state, n, sv_ref.lower(gates[:ns]))
    t_gates = (time.perf_counter() - t0) * len(gates) / ns
    t0 = time.perf_counter()
    sv_ref.tfim_energy_sv(state)
    t_energy = t_gates + time.perf_counter() - t0
    P = th.size
    out['sv_24q'] = {'seconds_per_gradient': 2 * P * t_energy, 'cores': 1, 'kind': 'port',
                     'sample': f'{ns} of {len(gates)} gates at 24 qubits (NumPy pair updates) '
                               f'and one TFIM energy timed, {t_energy:.1f} s per energy '
                               f'extrapolated to 2 x {P} central-difference evaluations'}
    n2 = 200
    th2 = rng.uniform(-math.pi, math.pi, 2 * n2 * layers)
    m = mps_ref.MPS(n2, 32, 1e-12)
    t0 = time.perf_counter()
    for name, qs, ps in sv_ref.hea_gates(n2, th2, layers):
        if len(qs) == 1:
            m.apply_1q(qs[0], mps_ref.gate_1q(name, ps))
        else:
            m.apply_adjacent_2q(qs[0], mps_ref.gate_2q(name, ps))
    mps_ref.tfim_energy(m)
    t_energy2 = time.perf_counter() - t0
    P2 = th2.size
    out['mps_200q_chi32'] = {'seconds_per_gradient': 2 * P2 * t_energy2, 'cores': 1,
                             'kind': 'port',
                             'sample': f'one energy of the 200-qubit chi_max 32 ansatz on the '
                                       f'MPS oracle (forward + TFIM energy), {t_energy2:.1f} s, '
                                       f'extrapolated to 2 x {P2} central-difference '
                                       f'evaluations'}
    return out


def traffic_from_profiles():
    """HBM bytes per pass launch from the committed rocprofv3 PMC summary, if present
    (the line `hbm_bytes_per_launch: N` of profiles/sv_pass_28q_traffic.md)."""
    path = os.path.join(ROOT, 'profiles', 'sv_pass_28q_traffic.md')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.strip().startswith('hbm_bytes_per_launch:'):
                return int(float(line.split(':', 1)[1].strip()))
    return None


def mps_leg():
    try:
        from qsim_amd import mps_bench
    except ImportError:
        return None
    return mps_bench.run()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus must be 1',
                          'n_gpus': args.gpus}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({'error': 'no GPU visible'}))
        sys.exit(2)

    if args.multi > 0:
        from qsim_amd import mps_bench
        print(json.dumps(mps_bench.run_multi(args.multi)))
        return
    if args.sv_sizes:
        for nq in (int(x) for x in args.sv_sizes.split(',')):
            print(json.dumps(sv_size(nq, args.steps, args.warmup)), flush=True)
        return
    sv = sv_leg(args.steps, args.warmup, args.calibration_copy,
                specialize=not args.interpreter)
    if args.pass_table:
        write_pass_table(args.pass_table, sv, 'python bench.py ' + ' '.join(sys.argv[1:]))
    value = sv['ngates'] * args.steps / (sv['gpu_ms'] / 1e3)
    achieved = sv['bytes_per_pass'] / sv['per_launch_s'] / 1e9
    out = {
        'metric': METRIC,
        'value': value,
        'unit': 'gates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': sv['gpu_ms'] / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'c128',
        'data': 'synthetic',
        'config': {'workload': '28q statevector: 20x[U3 all qubits + CZ ladder] + QFT-28 '
                               '(1520 gates), complex128, state resident in HBM',
                   'qubits': SV_QUBITS, 'gates': sv['ngates'], 'passes_per_step': sv['npass'],
                   'plan_ms': sv['plan_s'] * 1e3, 'plan_device_bytes': sv['plan_device_bytes'],
                   'pass_kernels': ('per-plan kernels (qs_sv_plan_specialize)'
                                    if sv['specialized'] else 'op-stream interpreter'),
                   'jit_ms': sv['jit_s'] * 1e3, 'jit_kernels': sv['jit_kernels'],
                   'hbm_gbps_whole_step': sv['npass'] * sv['bytes_per_pass'] * args.steps /
                   (sv['gpu_ms'] / 1e3) / 1e9},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBPS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBPS,
                     'traffic': traffic_from_profiles(),
                     'kernel': ('qs_sv_pass_* (per-plan pass kernels)' if sv['specialized']
                                else 'sv_pass_kernel<4>'),
                     'per_launch_us': sv['per_launch_s'] * 1e6,
                     'inplace_rmw_gbps': sv['inplace_rmw_gbps']},
        'check': {'norm_sq_after': sv['norm_sq']},
    }
    if not args.no_mps:
        m = mps_leg()
        if m is not None:
            out['mps'] = m
            from qsim_amd import mps_bench
            out['mps']['multi'] = mps_bench.run_multi(args.multi_sample, quiet=True)
            out['mps']['multi']['sample'] = (
                f'the first {args.multi_sample} circuits of the seeded 256-circuit list; the '
                f'full list: python bench.py --multi 256')
    if not args.no_mps:
        out['gradients'] = gradients_leg()
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_leg(args.cpu_gates)
        out['cpu_baseline']['cpu_model'] = cpu_model()
        out['cpu_baseline']['host_cpus'] = os.cpu_count()
        out['cpu_baseline']['openmp'] = cpu_omp_leg()
        if not args.no_mps and 'mps' in out:
            out['cpu_baseline']['mps'] = cpu_mps_leg()
            from threadpoolctl import threadpool_limits
            with threadpool_limits(limits=1):
                out['cpu_baseline']['gradients'] = cpu_gradients_leg()
    print(json.dumps(out))


if __name__ == '__main__':
    main()
