"""Drop-in for the reference's statevector driver (fpga_simulator_v07.py), MI355X edition.

Same public names, arguments, result dictionaries and exceptions as the
reference module; the pyxrt buffers and the HLS kernel are replaced by a
complex128 state in HBM and libqsim_hip.so's fused passes.

Deviations (each deliberate):
  * parameters are used in FP64 (the FPGA packs them as float32,
    fpga_simulator_v07.py:469-472), which removes a ~1e-7 error floor;
  * gates are applied as exact unitaries instead of the {1q, cx, cz} lowering;
    the lowering's global phase (exp(i*theta/4) per cp) is re-applied so the
    returned amplitudes equal the reference's;
  * mcx with more than two controls is an exact multi-controlled X (the
    reference asks Qiskit for a decomposition, which is not installed here);
  * no 512-gate chunking (any gate count goes to one planned launch
    sequence), and MAX_QUBITS_FPGA is the library's limit (34), not 28;
  * xclbin_path is accepted and ignored.
"""

from __future__ import annotations

import logging
import math
import os
import time
import uuid
from collections import defaultdict
from dataclasses import dataclass, field
from datetime import datetime
from typing import Dict, List, Optional, Tuple

import numpy as np

from qsim_amd import _lib
from qsim_amd.sv import StateVector, gates_to_records, plan_passes

logger = logging.getLogger(__name__)

VERSION = "07"
KERNEL_NAME = "quantum_fpga_kernel_complex128"
KERNEL_CLOCK_MHZ = 225.0          # kept for API compatibility (FPGA clock of the reference)
KERNEL_CLOCK_PERIOD_NS = 1e3 / KERNEL_CLOCK_MHZ
NUM_HBM_BANKS = 8                 # MI355X: 8 HBM3E stacks
MAX_QUBITS_FPGA = 34
GATE_WORDS = 8
KERNEL_MAX_GATES = 512            # the reference's chunk size; not a limit here
BYTES_PER_AMP = 16
BYTES_PER_F64 = 8

KERNEL_GATE_OPS: Dict[str, int] = {
    'id': 0, 'h': 1, 'x': 2, 'y': 3, 'z': 4, 's': 5, 't': 6, 'sdg': 7, 'tdg': 8,
    'sx': 9, 'sxdg': 10, 'rx': 11, 'ry': 12, 'rz': 13, 'p': 14, 'phase': 14, 'u1': 14,
    'u2': 15, 'u3': 16, 'u': 16, 'cx': 17, 'cnot': 17, 'cz': 18,
}
DECOMPOSABLE_GATES = frozenset({'mcx', 'mct'})
NON_UNITARY_GATES = frozenset({'measure', 'barrier', 'snapshot', 'delay', 'initialize'})
RESET_GATE = 'reset'
ALL_SUPPORTED_GATES: frozenset = frozenset({
    'id', 'h', 'x', 'y', 'z', 's', 't', 'sdg', 'tdg', 'sx', 'sxdg',
    'rx', 'ry', 'rz', 'p', 'phase', 'u1', 'u2', 'u3', 'u',
    'cx', 'cnot', 'cz', 'cy', 'ch', 'swap', 'cp', 'cphase',
    'crx', 'cry', 'crz', 'csx', 'dcx', 'iswap', 'ecr', 'rxx', 'ryy', 'rzz',
    'ccx', 'toffoli', 'ccnot', 'mcx', 'mct',
})
TRUSTED_KERNEL_GATES_DEFAULT = ALL_SUPPORTED_GATES
HBM_PEAK_GBPS = 8000.0


@dataclass
class GateProfile:
    gate_name: str; target_qubit: int; est_cycles: int; est_time_us: float
    hbm_read_kb: float; hbm_write_kb: float; est_energy_uj: float


@dataclass
class ValidationResult:
    passed: bool; norm: float; prob_sum: float; norm_ok: bool; prob_ok: bool
    expected_ok: Optional[bool] = None; expected_detail: str = ""
    norm_tolerance: float = 1e-6; prob_tolerance: float = 1e-6


@dataclass
class RunProfile:
    circuit_name: str = ""; num_qubits: int = 0; num_gates: int = 0
    num_chunks: int = 1; num_resets: int = 0
    host_distribute_us: float = 0.0; host_encode_us: float = 0.0
    xrt_kernel_us: float = 0.0; host_collect_us: float = 0.0
    host_reset_us: float = 0.0; host_measure_us: float = 0.0
    host_decompose_us: float = 0.0; total_wall_us: float = 0.0
    est_total_cycles: int = 0; est_kernel_time_us: float = 0.0
    total_hbm_read_mb: float = 0.0; total_hbm_write_mb: float = 0.0
    est_eff_bw_gbps: float = 0.0; total_energy_mj: float = 0.0
    avg_power_w: float = 0.0; energy_per_gate_uj: float = 0.0
    validation: Optional[ValidationResult] = None
    gate_profiles: List[GateProfile] = field(default_factory=list)
    gate_type_summary: Dict = field(default_factory=dict)
    num_passes: int = 0


class FPGAValidator:
    """State checks of the reference validator (fpga_simulator_v07.py:133-182)."""
    NORM_TOLERANCE = 1e-6
    PROB_TOLERANCE = 1e-6

    @staticmethod
    def check_norm(sv):
        n = float(np.linalg.norm(np.asarray(sv, dtype=np.complex128)))
        return abs(n - 1.0) <= FPGAValidator.NORM_TOLERANCE, n

    @staticmethod
    def check_probability_sum(sv):
        s = np.asarray(sv, dtype=np.complex128)
        ps = float(np.sum(s.real ** 2 + s.imag ** 2))
        return abs(ps - 1.0) <= FPGAValidator.PROB_TOLERANCE, ps

    @staticmethod
    def check_deterministic(sv, idx, tolerance=0.02, label=""):
        p = float(abs(sv[idx]) ** 2)
        ok = p >= 1.0 - tolerance
        return ok, f"{label}:P(|{idx}>)={p:.8f} ({'PASS' if ok else 'FAIL'})"

    @staticmethod
    def check_uniform_superposition(sv, nq, tolerance=0.05, label=""):
        exp = 1.0 / math.sqrt(2 ** nq)
        dev = float(np.max(np.abs(np.abs(np.asarray(sv, dtype=np.complex128)) - exp)))
        ok = dev <= tolerance
        return ok, f"{label}:max_dev={dev:.8f} ({'PASS' if ok else 'FAIL'})"

    @staticmethod
    def check_marginal_probability(sv, nq, pat, min_prob=0.5, label=""):
        idx = np.arange(2 ** nq, dtype=np.int64)
        mask = np.ones(idx.size, dtype=bool)
        for q, b in pat.items():
            mask &= ((idx >> q) & 1) == b
        s = np.asarray(sv, dtype=np.complex128)
        mp = float((s.real ** 2 + s.imag ** 2)[mask].sum())
        ok = mp >= min_prob
        return ok, f"{label}:P({pat})={mp:.8f} ({'PASS' if ok else 'FAIL'})"

    @classmethod
    def validate(cls, sv, nq, circuit_name="", det_idx=None, uniform_check=False,
                 marginal_pattern=None, marginal_min_prob=0.5, det_tolerance=0.02,
                 uniform_tolerance=0.05):
        no, n = cls.check_norm(sv)
        po, ps = cls.check_probability_sum(sv)
        eo, ed = None, ""
        if det_idx is not None:
            eo, ed = cls.check_deterministic(sv, det_idx, tolerance=det_tolerance,
                                             label=circuit_name)
        elif marginal_pattern is not None:
            eo, ed = cls.check_marginal_probability(sv, nq, marginal_pattern,
                                                    min_prob=marginal_min_prob,
                                                    label=circuit_name)
        elif uniform_check:
            eo, ed = cls.check_uniform_superposition(sv, nq, tolerance=uniform_tolerance,
                                                     label=circuit_name)
        return ValidationResult(passed=no and po and (eo is None or eo), norm=n, prob_sum=ps,
                                norm_ok=no, prob_ok=po, expected_ok=eo, expected_detail=ed,
                                norm_tolerance=cls.NORM_TOLERANCE,
                                prob_tolerance=cls.PROB_TOLERANCE)


class EnergyEstimator:
    """Kept for API compatibility; reports nothing for this hardware."""

    @staticmethod
    def gate_energy_uj(cyc, hbm):
        return 0.0

    @staticmethod
    def total_energy_mj(cyc, hbm):
        return 0.0, 0.0


def _validate_and_split_gates(gates):
    """Split gate tuples at resets (fpga_simulator_v07.py:330-369).

    Returns (segments, resets, meas) with resets as (qubit, segment index) and
    meas as (qubit, clbit) in encounter order; a missing clbit maps to the
    qubit's own index.
    """
    segments: List[list] = [[]]
    resets: List[Tuple[int, int]] = []
    meas: List[Tuple[int, int]] = []
    for gt in gates:
        if len(gt) == 2:
            nm, qs = gt
            ps = None
        elif len(gt) == 3:
            nm, qs, ps = gt
        else:
            raise ValueError(f"Bad gate tuple: {gt!r}")
        if not isinstance(nm, str):
            raise ValueError("Gate name must be str")
        name = nm.lower()
        if name in NON_UNITARY_GATES:
            if name == 'measure':
                cl = list(ps) if ps else []
                for i, q in enumerate(list(qs or [])):
                    meas.append((int(q), int(cl[i]) if i < len(cl) else int(q)))
            continue
        if name == RESET_GATE:
            for q in (qs or []):
                resets.append((int(q), len(segments) - 1))
            segments.append([])
            continue
        if name not in ALL_SUPPORTED_GATES:
            raise ValueError(f"Unsupported gate '{nm}'")
        if name in DECOMPOSABLE_GATES:
            qs = [int(q) for q in qs]
            if not qs:
                raise ValueError("mcx needs at least one qubit")
            if qs[-1] in qs[:-1]:
                raise ValueError(f"mcx target {qs[-1]} in controls {qs[:-1]}")
        segments[-1].append((name, list(qs), ps))
    if segments and not segments[-1]:
        segments.pop()
    return segments, resets, meas


class QuantumFPGACircuit:
    """Statevector engine with the reference's interface (fpga_simulator_v07.py:374-641)."""

    def __init__(self, xclbin_path=None, device_index=0):
        _lib.ensure_device(device_index)
        self.xclbin_path = xclbin_path
        self.device_index = device_index
        self._sv: Optional[StateVector] = None
        self._debug = bool(int(os.environ.get("FPGA_SIM_DEBUG", "0") or "0"))
        logger.info("QuantumFPGACircuit v%s (complex128, MI355X HIP) ready", VERSION)

    # -- device state -------------------------------------------------------
    def _ensure_hbm(self, nq):
        if self._sv is None or self._sv.nq != nq:
            self._sv = None
            self._sv = StateVector(nq, self.device_index, init_zero=False)
        return self._sv

    def _distribute_statevector(self, nq):
        self._ensure_hbm(nq).set_zero()

    def _distribute_arbitrary_state(self, sv, nq):
        self._ensure_hbm(nq).load(sv)

    def _collect_statevector(self, nq):
        return self._ensure_hbm(nq).to_numpy()

    def reset_state(self, nq):
        self._distribute_statevector(nq)

    def probe_native_gates(self, verbose=False):
        if verbose:
            print(f"  v{VERSION} native: {sorted(ALL_SUPPORTED_GATES)}")
        return frozenset(ALL_SUPPORTED_GATES)

    @property
    def state_tensor(self):
        """The device state (torch complex128 tensor) of the last run."""
        return None if self._sv is None else self._sv.tensor

    # -- execution ----------------------------------------------------------
    def run_circuit(self, num_qubits, gates, shots=1024, validate=True,
                    det_idx=None, uniform_check=False, marginal_pattern=None,
                    marginal_min_prob=0.5, det_tolerance=0.02, uniform_tolerance=0.05,
                    circuit_name="", skip_init=False, num_clbits=None):
        """Run a circuit; same arguments and result dict as the reference."""
        import torch
        if not (1 <= num_qubits <= MAX_QUBITS_FPGA):
            raise ValueError(f"num_qubits must be in [1,{MAX_QUBITS_FPGA}]")
        t0 = time.perf_counter_ns()
        N = 2 ** num_qubits

        tlo = time.perf_counter_ns()
        segments, resets, meas = _validate_and_split_gates(gates)
        seg_records = []
        phase = 1.0 + 0.0j
        for seg in segments:
            recs, ph = gates_to_records(seg, num_qubits)
            seg_records.append(recs)
            phase *= ph
        hdu = (time.perf_counter_ns() - tlo) / 1e3

        sv_dev = self._ensure_hbm(num_qubits)
        if not skip_init:
            sv_dev.set_zero()
        torch.cuda.synchronize(self.device_index)
        td = time.perf_counter_ns()

        reset_by_seg = defaultdict(list)
        for q, si in resets:
            reset_by_seg[si].append(q)
        kernel_us = 0.0
        reset_us = 0.0
        ng = 0
        npass = 0
        for si, recs in enumerate(seg_records):
            tk0 = time.perf_counter_ns()
            sv_dev.apply_records(recs)
            torch.cuda.synchronize(self.device_index)
            kernel_us += (time.perf_counter_ns() - tk0) / 1e3
            ng += len(segments[si])
            if recs:
                npass += plan_passes(num_qubits, recs)
            for q in reset_by_seg.get(si, []):
                tr0 = time.perf_counter_ns()
                sv_dev.reset_qubit(q)
                reset_us += (time.perf_counter_ns() - tr0) / 1e3
        # resets recorded after the last non-empty segment
        for si in sorted(reset_by_seg):
            if si >= len(seg_records):
                for q in reset_by_seg[si]:
                    tr0 = time.perf_counter_ns()
                    sv_dev.reset_qubit(q)
                    reset_us += (time.perf_counter_ns() - tr0) / 1e3
        tk = time.perf_counter_ns()

        # normalise on the device, re-apply the lowering phase, then collect
        nb = math.sqrt(sv_dev.norm_sq())
        scale = phase / nb if nb > 1e-12 else phase
        if scale != 1.0:
            sv_dev.scale(scale)
        sv = sv_dev.to_numpy()
        tc = time.perf_counter_ns()

        val = None
        if validate:
            val = FPGAValidator.validate(sv=sv, nq=num_qubits, circuit_name=circuit_name,
                                         det_idx=det_idx, uniform_check=uniform_check,
                                         marginal_pattern=marginal_pattern,
                                         marginal_min_prob=marginal_min_prob,
                                         det_tolerance=det_tolerance,
                                         uniform_tolerance=uniform_tolerance)
            if not val.passed:
                logger.warning(f"[{circuit_name}] validation FAILED norm={val.norm:.10f}")

        nc = num_qubits if num_clbits is None else int(num_clbits)
        if nc <= 0:
            nc = num_qubits
        meas_pairs = meas if meas else [(q, q) for q in range(min(num_qubits, nc))]
        counts: Dict[str, int] = {}
        memory: List[str] = []
        if shots > 0:
            # np.random.choice(N, shots, p=|sv|^2 / sum) draws `shots` uniforms from the
            # global generator and returns the first index whose cumulative probability
            # exceeds each; the same draws go to the device sampler.
            if not nb > 0.0:
                raise ValueError('probabilities do not sum to 1')   # as np.random.choice
            u = np.random.random_sample(shots)
            indices = sv_dev.sample_indices(u)
            for idx in indices:
                qb = format(int(idx), f'0{num_qubits}b')
                cb = ['0'] * nc
                for q, c in meas_pairs:
                    if 0 <= q < num_qubits and 0 <= c < nc:
                        cb[nc - 1 - c] = qb[num_qubits - 1 - q]
                key = ''.join(cb)
                counts[key] = counts.get(key, 0) + 1
                memory.append(key)
        te = time.perf_counter_ns()
        measured_qubits = sorted({q for q, _ in meas_pairs})

        hbm_bytes = npass * 2 * BYTES_PER_AMP * N
        prof = RunProfile(
            circuit_name=circuit_name, num_qubits=num_qubits, num_gates=ng,
            num_chunks=max(1, len(seg_records)), num_resets=len(resets),
            host_distribute_us=(td - t0) / 1e3, host_encode_us=0.0, xrt_kernel_us=kernel_us,
            host_collect_us=(tc - tk) / 1e3, host_reset_us=reset_us,
            host_measure_us=(te - tc) / 1e3, host_decompose_us=hdu,
            total_wall_us=(te - t0) / 1e3, est_total_cycles=0, est_kernel_time_us=kernel_us,
            total_hbm_read_mb=hbm_bytes / 2 / 2 ** 20, total_hbm_write_mb=hbm_bytes / 2 / 2 ** 20,
            est_eff_bw_gbps=(hbm_bytes / 1e9) / (kernel_us / 1e6) if kernel_us > 0 else 0.0,
            validation=val, num_passes=npass)
        return {'statevector': sv.copy(), 'counts': counts, 'memory': memory,
                'measured_qubits': measured_qubits, 'clbit_map': meas_pairs, 'num_clbits': nc,
                'success': True, 'execution_time': prof.total_wall_us / 1e6,
                'timing': {'distribute_us': prof.host_distribute_us,
                           'decompose_us': prof.host_decompose_us,
                           'kernel_us': prof.xrt_kernel_us,
                           'collect_us': prof.host_collect_us,
                           'reset_us': prof.host_reset_us,
                           'measure_us': prof.host_measure_us,
                           'total_us': prof.total_wall_us,
                           'num_chunks': prof.num_chunks,
                           'num_resets': prof.num_resets},
                'profile': prof, 'validation': val}


class FPGASimulator:
    """Qiskit-style backend over QuantumFPGACircuit (fpga_simulator_v07.py:646-738)."""
    _DEFAULT_XCLBIN = "quantum_fpga_kernel_complex128.xclbin"

    def __init__(self, xclbin_path=None, device_index=0, **_):
        self._fpga = QuantumFPGACircuit(xclbin_path or self._DEFAULT_XCLBIN, device_index)
        self.name = f"fpga_simulator_v{VERSION}"
        self._last_profile = None
        self._configuration = None

    def configuration(self):
        return self._configuration

    @property
    def last_profile(self):
        return self._last_profile

    def run(self, circuits, shots=1024, seed_simulator=None, **kwargs):
        t0 = time.time()
        if not isinstance(circuits, list):
            circuits = [circuits]
        if seed_simulator is not None:
            np.random.seed(seed_simulator)
        results = []
        for circuit in circuits:
            gates = []
            has_meas = False
            for inst in circuit.data:
                op = inst.operation.name
                try:
                    qs = [circuit.find_bit(q).index for q in inst.qubits]
                except Exception:
                    qs = [circuit.qubits.index(q) for q in inst.qubits]
                if op.lower() == 'measure':
                    has_meas = True
                    try:
                        cs = [circuit.find_bit(c).index for c in inst.clbits]
                    except Exception:
                        cs = [circuit.clbits.index(c) for c in inst.clbits]
                    gates.append((op, qs, cs))
                    continue
                gates.append((op, qs, getattr(inst.operation, 'params', None)))
            if not has_meas and shots > 0:
                ql = list(range(circuit.num_qubits))
                gates.append(('measure', ql, ql))
            cname = getattr(circuit, 'name', 'unnamed')
            nc = circuit.num_clbits if circuit.num_clbits > 0 else circuit.num_qubits
            cr = self._fpga.run_circuit(circuit.num_qubits, gates, shots=shots,
                                        circuit_name=cname, num_clbits=nc)
            self._last_profile = cr['profile']
            sv = cr.get('statevector')
            pf = cr['profile']
            results.append({
                'success': True, 'shots': shots,
                'data': {'counts': cr['counts'],
                         'statevector': sv.tolist() if isinstance(sv, np.ndarray) else None,
                         'execution_time': cr['execution_time']},
                'header': {'name': pf.circuit_name,
                           'metadata': {'fpga_enabled': False, 'gpu_enabled': True,
                                        'precision': 'complex128',
                                        'kernel_name': KERNEL_NAME,
                                        'hbm_banks': NUM_HBM_BANKS,
                                        'execution_mode': 'hip-gfx950-complex128-fused',
                                        'xrt_kernel_us': pf.xrt_kernel_us,
                                        'num_chunks': pf.num_chunks,
                                        'num_resets': pf.num_resets,
                                        'num_passes': pf.num_passes,
                                        'eff_bw_gbps': pf.est_eff_bw_gbps,
                                        'num_clbits': cr['num_clbits'],
                                        'clbit_map': cr['clbit_map'],
                                        'norm': pf.validation.norm if pf.validation else None,
                                        'validation_passed':
                                            pf.validation.passed if pf.validation else None},
                           'timing': cr['timing'], 'memory_slots': cr['num_clbits'],
                           'qreg_sizes': [['q', circuit.num_qubits]],
                           'qubit_labels': [['q', i] for i in range(circuit.num_qubits)],
                           'creg_sizes': [['c', cr['num_clbits']]],
                           'clbit_labels': [['c', i] for i in range(cr['num_clbits'])]}})
        rd = {'backend_name': self.name, 'backend_version': VERSION, 'qobj_id': None,
              'job_id': kwargs.get('job_id', str(uuid.uuid4())),
              'success': all(r['success'] for r in results), 'results': results,
              'date': datetime.now().isoformat(), 'status': 'COMPLETED', 'header': {},
              'metadata': {'shots': shots, 'seed_simulator': seed_simulator,
                           'precision': 'complex128', 'hbm_banks': NUM_HBM_BANKS,
                           'fpga_enabled': False, 'total_run_time': time.time() - t0}}
        return _Result(rd)


class _Result:
    """Minimal result object (the reference's fallback when Qiskit is absent)."""

    def __init__(self, d):
        self._d = d

    def get_counts(self, idx=0):
        return self._d['results'][idx if isinstance(idx, int) else 0]['data']['counts']

    def get_statevector(self, idx=0):
        return np.asarray(self._d['results'][idx]['data']['statevector'], dtype=np.complex128)

    def result(self):
        return self

    def to_dict(self):
        return self._d


__version__ = VERSION
__all__ = ['VERSION', 'KERNEL_NAME', 'KERNEL_GATE_OPS', 'ALL_SUPPORTED_GATES',
           'NUM_HBM_BANKS', 'MAX_QUBITS_FPGA', 'GATE_WORDS', 'KERNEL_MAX_GATES', 'BYTES_PER_AMP',
           'GateProfile', 'ValidationResult', 'RunProfile', 'FPGAValidator', 'EnergyEstimator',
           'QuantumFPGACircuit', 'FPGASimulator']
