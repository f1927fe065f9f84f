"""MPS engine over libqsim_hip.so.

A chain of n sites lives in HBM as one complex128 tensor of n slots, each slot
sized for a (chi_max, 2, chi_max) tensor; the actual bond dimensions stay on
the device (int32 bond[n+1]), so consecutive two-site updates never wait for
the host.  Two-site updates on disjoint bonds are submitted as one batch
(one contraction launch + one SVD launch for the whole batch).

Semantics follow mps_core.MPSState (mps_core.py:172-494): |0..0> start,
sqrt(S) absorbed into both neighbours, norm_log bookkeeping of _svd_truncate.
"""

from __future__ import annotations

import ctypes
import math
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

PRECISIONS = {'fp64': _lib.QS_PREC_FP64, 'fp32': _lib.QS_PREC_FP32_MFMA,
              'fp32_mfma': _lib.QS_PREC_FP32_MFMA}


def max_chi() -> int:
    return _lib.lib().qs_mps_max_chi()


class MPSChain:
    """Device-resident MPS of n qubits with bond dimensions up to chi_max."""

    def __init__(self, n: int, chi_max: int, cutoff: float = 1e-12, device: int = 0,
                 precision: str = 'fp64'):
        import torch
        _lib.ensure_device(device)
        if n < 1:
            raise ValueError('n must be >= 1')
        if not 1 <= chi_max <= max_chi():
            raise ValueError(f'chi_max must be in [1, {max_chi()}], got {chi_max}')
        if precision not in PRECISIONS:
            raise ValueError(f'precision must be one of {sorted(PRECISIONS)}')
        self.n = n
        self.chi_max = chi_max
        self.cutoff = float(cutoff)
        self.device = device
        self.precision = precision
        self.stride = 2 * chi_max * chi_max
        dev = f'cuda:{device}'
        self.sites = torch.zeros(n * self.stride, dtype=torch.complex128, device=dev)
        self.sites.view(n, self.stride)[:, 0] = 1.0
        self.bond = torch.ones(n + 1, dtype=torch.int32, device=dev)
        self._dlogs: List = []          # device tensors of per-update norm_log increments
        self._ws = None
        self._ws_tasks = 0
        self._status = None
        self._err = None
        self._last_batch = 0
        self.updates = 0

    # -- plumbing ------------------------------------------------------------
    def _stream(self):
        return _lib.current_stream(self.device)

    def _workspace(self, ntasks: int):
        import torch
        if self._ws is None or self._ws_tasks < ntasks:
            nbytes = _lib.lib().qs_mps_workspace_bytes(ntasks, self.chi_max)
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.sites.device)
            self._ws_tasks = ntasks
            self._status = torch.zeros(ntasks, dtype=torch.int32, device=self.sites.device)
        return self._ws, self._status

    def _status_buffer(self, ntasks: int):
        import torch
        if self._status is None or self._status.numel() < ntasks:
            self._status = torch.zeros(ntasks, dtype=torch.int32, device=self.sites.device)
            self._ws = None
            self._ws_tasks = 0
        return self._status

    # -- gates ---------------------------------------------------------------
    def apply_2q_batch(self, updates: Sequence[Tuple[int, np.ndarray]],
                       precision: Optional[str] = None, allocation_free: bool = False) -> None:
        """Two-site updates on bonds (i, i+1), all disjoint, in one launch sequence.

        allocation_free: run through qs_mps_update_batch_static, whose workspace is the
        library's static device scratch (no workspace allocated here; batches beyond its
        size are refused)."""
        import torch
        k = len(updates)
        if k == 0:
            return
        prec = PRECISIONS[precision or self.precision]
        dl = torch.zeros(k, dtype=torch.float64, device=self.sites.device)
        tasks = (_lib.QsMpsTask * k)()
        base = self.sites.data_ptr()
        for j, (i, g) in enumerate(updates):
            if not 0 <= i < self.n - 1:
                raise ValueError(f'bond ({i},{i + 1}) outside the chain of {self.n} sites')
            t = tasks[j]
            t.sites = base
            t.bond = self.bond.data_ptr()
            t.dlog = dl.data_ptr() + 8 * j
            t.site_stride = self.stride
            t.site = int(i)
            t.n_sites = self.n
            t.chi_max = self.chi_max
            t.cutoff = self.cutoff
            gf = np.asarray(g, dtype=np.complex128).reshape(16)
            vals = np.empty(32)
            vals[0::2] = gf.real
            vals[1::2] = gf.imag
            t.gate[:] = vals.tolist()
        if allocation_free:
            status = self._status_buffer(k)
            _lib.call('qs_mps_update_batch_static', tasks, k, prec,
                      ctypes.c_void_p(status.data_ptr()), self._stream())
        else:
            ws, status = self._workspace(k)
            _lib.call('qs_mps_update_batch', tasks, k, prec, ctypes.c_void_p(ws.data_ptr()),
                      ws.numel(), ctypes.c_void_p(status.data_ptr()), self._stream())
        self._dlogs.append(dl)
        self._last_batch = k
        # running minimum of the statuses (negative = failure), kept on the device
        m = status[:k].min().reshape(1)
        self._err = m if self._err is None else torch.minimum(self._err, m)
        self.updates += k

    def apply_2q(self, i: int, g: np.ndarray) -> None:
        self.apply_2q_batch([(i, g)])

    def apply_1q(self, site: int, u: np.ndarray) -> None:
        if not 0 <= site < self.n:
            raise ValueError(f'site {site} outside the chain')
        m = np.asarray(u, dtype=np.complex128).reshape(4)
        arr = (ctypes.c_double * 8)(*[float(v) for pair in zip(m.real, m.imag) for v in pair])
        _lib.call('qs_mps_apply_1q', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, int(site), arr,
                  self._stream())

    def apply_1q_batch(self, gates: Sequence[Tuple[int, np.ndarray]]) -> None:
        """One-site gates on distinct sites, 32 per launch (qs_mps_apply_1q_batch)."""
        k = len(gates)
        if k == 0:
            return
        sites = (ctypes.c_int32 * k)(*[int(s) for s, _ in gates])
        vals = np.empty((k, 8))
        for j, (_, u) in enumerate(gates):
            m = np.asarray(u, dtype=np.complex128).reshape(4)
            vals[j, 0::2] = m.real
            vals[j, 1::2] = m.imag
        vals = np.ascontiguousarray(vals)
        _lib.call('qs_mps_apply_1q_batch', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, self.n, sites,
                  vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), k, self._stream())

    def reset_site(self, site: int) -> None:
        """Zero the |1> half of a site and renormalise it, or set it to |0>
        (fpga_mps_simulator.py:530-552), on the device."""
        _lib.call('qs_mps_reset_site', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, self.n, int(site),
                  self._stream())

    def check_status(self) -> None:
        """Raise if any two-site update so far failed (status < 0)."""
        if self._err is not None:
            worst = int(self._err.cpu().item())
            if worst < 0:
                raise RuntimeError(f'MPS update status {worst} '
                                   f'(QS_E_NOCONV={_lib.QS_E_NOCONV}, QS_E_CHI={_lib.QS_E_CHI})')

    def last_sweeps(self) -> np.ndarray:
        """Jacobi sweeps of each update of the last batch."""
        return self._status.cpu().numpy()[:self._last_batch].copy()

    # -- host views -----------------------------------------------------------
    @property
    def norm_log(self) -> float:
        total = 0.0
        for d in self._dlogs:
            for v in d.cpu().numpy().tolist():
                total += v
        return total

    def bond_dims(self) -> List[int]:
        return self.bond.cpu().numpy()[1:self.n].astype(int).tolist()

    def max_bond(self) -> int:
        return 1 if self.n <= 1 else max(self.bond_dims())

    def site_tensors(self) -> List[np.ndarray]:
        b = self.bond.cpu().numpy()
        flat = self.sites.view(self.n, self.stride).cpu().numpy()
        return [flat[i, :b[i] * 2 * b[i + 1]].reshape(b[i], 2, b[i + 1]).copy()
                for i in range(self.n)]

    def set_site(self, i: int, t: np.ndarray) -> None:
        """Overwrite site i with a (chi_l, 2, chi_r) tensor whose dims match the bonds."""
        import torch
        b = self.bond.cpu().numpy()
        t = np.ascontiguousarray(np.asarray(t, dtype=np.complex128))
        if t.shape != (b[i], 2, b[i + 1]):
            raise ValueError(f'site {i}: shape {t.shape} != ({b[i]}, 2, {b[i + 1]})')
        row = self.sites.view(self.n, self.stride)[i]
        row[:t.size].copy_(torch.from_numpy(t.reshape(-1)))

    def statevector(self) -> np.ndarray:
        """Dense state with qubit q <-> index bit q, times exp(norm_log), contracted on
        the device (mps_core.py:433-461; the reference refuses n > 24)."""
        import torch
        if self.n > 24:
            raise ValueError(f"Statevector reconstruction for n={self.n} > 24 not supported.")
        dev = self.sites.device
        out = torch.empty(1 << self.n, dtype=torch.complex128, device=dev)
        work = torch.empty(1 << self.n, dtype=torch.complex128, device=dev)
        _lib.call('qs_mps_statevector', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, self.n, self.chi_max,
                  float(self.norm_log), ctypes.c_void_p(out.data_ptr()),
                  ctypes.c_void_p(work.data_ptr()), out.numel() * 16, self._stream())
        return out.cpu().numpy()

    def to_host(self) -> 'HostMPS':
        """Host copy of the chain; it keeps a handle on this device chain, so its
        read-outs run on the device until the copy is modified."""
        h = HostMPS(self.site_tensors(), self.norm_log, self.chi_max, self.cutoff)
        h.device_chain = self
        return h

    @classmethod
    def from_host(cls, host: 'HostMPS', device: int = 0, precision: str = 'fp64') -> 'MPSChain':
        """A device chain holding a host chain's site tensors (norm_log is not carried)."""
        import torch
        chi = max([1] + [t.shape[2] for t in host.sites[:-1]])
        ch = cls(host.n, max(chi, min(host.chi_max, max_chi())), host.svd_cutoff, device,
                 precision)
        b = np.array([1] + [t.shape[2] for t in host.sites[:-1]] + [1], dtype=np.int32)
        ch.bond.copy_(torch.from_numpy(b))
        for i, t in enumerate(host.sites):
            ch.set_site(i, t)
        return ch

    # -- shot sampling ----------------------------------------------------------
    def sample(self, shots: int, measured: Optional[List[int]] = None,
               rng: Optional[np.random.Generator] = None,
               max_shots_per_launch: int = 65536) -> List[str]:
        """Sequential conditional sampling on the device (mps_core.py:371-420).

        Bitstrings as the reference returns them: qubit n-1 first; with a measured
        subset, one character per measured site q in descending order holding the
        outcome of site n-1-q, the mirrored site (mps_core.py:415-417).  The uniforms
        are drawn from `rng` in the reference's order (one per site per shot,
        shot-major), so a seeded generator gives the reference's samples.
        """
        import torch
        rng = rng if rng is not None else np.random.default_rng()
        n = self.n
        measured = list(range(n)) if measured is None else [int(q) for q in measured]
        for q in measured:
            if not 0 <= q < n:
                raise ValueError(f'measured qubit {q} outside the chain of {n}')
        dev = self.sites.device
        lib = _lib.lib()
        nbytes = lib.qs_mps_env_bytes(n, self.chi_max)
        envs = torch.empty(nbytes // 16, dtype=torch.complex128, device=dev)
        stream = self._stream()
        _lib.call('qs_mps_right_envs', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, n, self.chi_max,
                  ctypes.c_void_p(envs.data_ptr()), nbytes, stream)
        cols = (list(range(n - 1, -1, -1)) if set(measured) == set(range(n))
                else [n - 1 - q for q in sorted(measured, reverse=True)])
        out: List[str] = []
        for b0 in range(0, shots, max_shots_per_launch):
            k = min(max_shots_per_launch, shots - b0)
            u = torch.from_numpy(rng.random(k * n)).to(dev)
            bits = torch.empty(k * n, dtype=torch.uint8, device=dev)
            _lib.call('qs_mps_sample', ctypes.c_void_p(self.sites.data_ptr()),
                      ctypes.c_void_p(self.bond.data_ptr()), self.stride, n, self.chi_max,
                      ctypes.c_void_p(envs.data_ptr()), ctypes.c_void_p(u.data_ptr()), k,
                      ctypes.c_void_p(bits.data_ptr()), stream)
            b = bits.cpu().numpy().reshape(k, n)[:, cols] + ord('0')
            out.extend(row.tobytes().decode() for row in b)
        return out

    def get_counts(self, shots: int, measured: Optional[List[int]] = None,
                   rng: Optional[np.random.Generator] = None) -> Dict[str, int]:
        c: Dict[str, int] = {}
        for s in self.sample(shots, measured, rng):
            c[s] = c.get(s, 0) + 1
        return c

    # -- transfer-matrix read-outs on the device ---------------------------------
    def _left_envs(self):
        import torch
        nbytes = _lib.lib().qs_mps_env_bytes(self.n, self.chi_max)
        left = torch.empty(nbytes // 16, dtype=torch.complex128, device=self.sites.device)
        logs = torch.zeros(self.n, dtype=torch.float64, device=self.sites.device)
        _lib.call('qs_mps_left_envs', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, self.n, self.chi_max,
                  ctypes.c_void_p(left.data_ptr()), nbytes, ctypes.c_void_p(logs.data_ptr()),
                  self._stream())
        return left, logs

    def environments(self):
        """(left, right) environments of <psi|psi> on the device, each trace-normalised
        (qs_mps_left_envs / qs_mps_right_envs)."""
        import torch
        nbytes = _lib.lib().qs_mps_env_bytes(self.n, self.chi_max)
        left, _ = self._left_envs()
        right = torch.empty(nbytes // 16, dtype=torch.complex128, device=self.sites.device)
        _lib.call('qs_mps_right_envs', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, self.n, self.chi_max,
                  ctypes.c_void_p(right.data_ptr()), nbytes, self._stream())
        return left, right

    def entanglement_entropies(self, bonds: Optional[Sequence[int]] = None) -> np.ndarray:
        """Entanglement entropy of each bond (all by default) by the reference's definition
        (mps_core.py:465-483), computed on the device from the left Gram environments
        (qs_mps_entanglement_entropy)."""
        import torch
        bonds = list(range(self.n - 1)) if bonds is None else [int(b) for b in bonds]
        for b in bonds:
            if not 0 <= b < self.n - 1:
                raise ValueError(f"bond {b} out of range [0, {self.n - 2}]")
        if not bonds:
            return np.zeros(0)
        left, logs = self._left_envs()
        out = torch.zeros(len(bonds), dtype=torch.float64, device=self.sites.device)
        nbytes = _lib.lib().qs_mps_entropy_work_bytes(len(bonds), self.chi_max)
        work = torch.empty(nbytes, dtype=torch.uint8, device=self.sites.device)
        arr = (ctypes.c_int32 * len(bonds))(*bonds)
        _lib.call('qs_mps_entanglement_entropy', ctypes.c_void_p(self.bond.data_ptr()), self.n,
                  self.chi_max, ctypes.c_void_p(left.data_ptr()),
                  ctypes.c_void_p(logs.data_ptr()), arr, len(bonds),
                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(work.data_ptr()), nbytes,
                  self._stream())
        return out.cpu().numpy()

    def entanglement_entropy(self, bond: int) -> float:
        return float(self.entanglement_entropies([bond])[0])

    def bond_spectrum(self, bond: int) -> np.ndarray:
        """The reference's bond_spectrum (mps_core.py:485-494): singular values of site
        `bond` after sites 0..bond are left-normalised by QR, that site included.  A
        left-normalised site is an isometry, so these are ones; their number is
        min(2 chi'_l, chi_r) with chi'_l the bond dimension the earlier QRs leave
        (chi'_0 = 1, chi'_{i+1} = min(2 chi'_i, chi_r(i))), read from the device bond
        dimensions."""
        if not 0 <= bond < self.n:
            raise ValueError(f'bond {bond} outside the chain')
        b = self.bond.cpu().numpy()
        cl = 1
        for i in range(bond):
            cl = min(2 * cl, int(b[i + 1]))
        return np.ones(min(2 * cl, int(b[bond + 1])))

    def _expect(self, op_a, op_b, span: int) -> np.ndarray:
        import torch
        dev = self.sites.device
        left, right = self.environments()
        cnt = self.n - span + 1
        out = torch.zeros(max(cnt, 1), dtype=torch.float64, device=dev)
        work = torch.empty(_lib.lib().qs_mps_expect_work_bytes(self.chi_max), dtype=torch.uint8,
                           device=dev)

        def mat(o):
            m = np.asarray(o, dtype=np.complex128).reshape(4)
            return (ctypes.c_double * 8)(*[float(v) for z in m for v in (z.real, z.imag)])
        _lib.call('qs_mps_expect', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride, self.n, self.chi_max,
                  ctypes.c_void_p(left.data_ptr()), ctypes.c_void_p(right.data_ptr()),
                  mat(op_a), mat(op_b) if op_b is not None else None, span,
                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(work.data_ptr()),
                  work.numel(), self._stream())
        return out.cpu().numpy()[:max(cnt, 0)]

    def expectations(self, op) -> np.ndarray:
        """<O_i> / <psi|psi> for a 2x2 operator on every site, on the device."""
        return self._expect(op, None, 1)

    def pair_expectations(self, op_a, op_b) -> np.ndarray:
        """<O_i O'_{i+1}> / <psi|psi> for i = 0..n-2, on the device."""
        return self._expect(op_a, op_b, 2)

    def z_expectations(self) -> np.ndarray:
        return self.expectations(np.diag([1.0, -1.0]))

    def zz_expectations(self) -> np.ndarray:
        z = np.diag([1.0, -1.0])
        return self.pair_expectations(z, z)

    def overlap_log(self, other: 'MPSChain'):
        """(log |<self|other>|, arg <self|other>) with both norm_log factors, by transfer
        matrices on the device (qs_mps_overlap)."""
        import torch
        if other.n != self.n or other.chi_max != self.chi_max:
            raise ValueError('overlap needs chains of equal length and chi_max')
        chi = self.chi_max
        nbytes = _lib.lib().qs_mps_overlap_work_bytes(self.n, chi)
        work = torch.empty(nbytes, dtype=torch.uint8, device=self.sites.device)
        out = (ctypes.c_double * 2)()
        _lib.call('qs_mps_overlap', ctypes.c_void_p(self.sites.data_ptr()),
                  ctypes.c_void_p(self.bond.data_ptr()), self.stride,
                  ctypes.c_void_p(other.sites.data_ptr()), ctypes.c_void_p(other.bond.data_ptr()),
                  other.stride, self.n, chi, ctypes.c_void_p(work.data_ptr()), nbytes, out,
                  self._stream())
        return out[0] + self.norm_log + other.norm_log, out[1]

    def fidelity(self, other: 'MPSChain') -> float:
        """|<a|b>|^2 / (<a|a> <b|b>), on the device."""
        ab, _ = self.overlap_log(other)
        aa, _ = self.overlap_log(self)
        bb, _ = other.overlap_log(other)
        return float(math.exp(2 * ab - aa - bb))


class HostMPS:
    """Host copy of a chain with the interface of mps_core.MPSState (mps_core.py:172-494):
    the object FPGAMPSResult.get_mps returns.  Site i is a (chi_l, 2, chi_r) complex128
    array; the state is exp(norm_log) times the chain's contraction.  The canonical-form
    helpers and read-outs follow the reference's definitions; shots are drawn on the
    device (MPSChain.sample)."""

    def __init__(self, sites: List[np.ndarray], norm_log: float, chi_max: Optional[int] = None,
                 svd_cutoff: float = 1e-12):
        self.sites = sites
        self.n = len(sites)
        self.norm_log = norm_log
        self.chi_max = chi_max if chi_max is not None else max([1] + [t.shape[2] for t in sites])
        self.svd_cutoff = svd_cutoff
        # the device chain this copy was read from (MPSChain.to_host); cleared by any
        # change to the copy, after which read-outs run on the host tensors
        self.device_chain = None

    def copy(self) -> 'HostMPS':
        return HostMPS([t.copy() for t in self.sites], self.norm_log, self.chi_max,
                       self.svd_cutoff)

    def chi_l(self, i: int) -> int:
        return self.sites[i].shape[0]

    def chi_r(self, i: int) -> int:
        return self.sites[i].shape[2]

    def bond_dims(self) -> List[int]:
        return [t.shape[2] for t in self.sites[:-1]]

    def max_bond(self) -> int:
        return 1 if self.n <= 1 else max(self.bond_dims())

    # -- canonical forms (mps_core.py:221-263) ----------------------------------
    def left_normalize_site(self, i: int) -> None:
        """Site i as a (2 chi_l) x chi_r matrix = Q R (reduced QR): site i becomes Q and R
        is absorbed into site i + 1 (mps_core.py:221-234)."""
        self.device_chain = None
        a = self.sites[i]
        cl, d, cr = a.shape
        q, r = np.linalg.qr(a.reshape(cl * d, cr))
        self.sites[i] = q.reshape(cl, d, q.shape[1])
        if i + 1 < self.n:
            self.sites[i + 1] = np.tensordot(r, self.sites[i + 1], axes=([1], [0]))

    def right_normalize_site(self, i: int) -> None:
        """Site i as a chi_l x (2 chi_r) matrix = L Q (QR of its adjoint): site i becomes Q
        and L is absorbed into site i - 1 (mps_core.py:236-251)."""
        self.device_chain = None
        a = self.sites[i]
        cl, d, cr = a.shape
        q, r = np.linalg.qr(a.reshape(cl, d * cr).conj().T)
        self.sites[i] = q.conj().T.reshape(q.shape[1], d, cr)
        if i >= 1:
            self.sites[i - 1] = np.tensordot(self.sites[i - 1], r.conj().T, axes=([2], [0]))

    def left_canonicalize(self) -> None:
        for i in range(self.n - 1):
            self.left_normalize_site(i)

    def right_canonicalize(self) -> None:
        for i in range(self.n - 1, 0, -1):
            self.right_normalize_site(i)

    def apply_1q(self, site: int, gate: np.ndarray) -> None:
        """A'[l, s', r] = sum_s G[s', s] A[l, s, r] (mps_core.py:267-275)."""
        self.device_chain = None
        self.sites[site] = np.einsum('ts,lsr->ltr', np.asarray(gate, dtype=np.complex128),
                                     self.sites[site])

    def get_statevector(self) -> np.ndarray:
        """Dense state with qubit q <-> index bit q (mps_core.py:433-461)."""
        if self.n > 24:
            raise ValueError(f"Statevector reconstruction for n={self.n} > 24 not supported.")
        acc = self.sites[0][0]
        for t in self.sites[1:]:
            acc = np.tensordot(acc, t, axes=([1], [0])).reshape(-1, t.shape[2])
        sv = acc[:, 0] * math.exp(self.norm_log)
        return sv.reshape([2] * self.n).transpose(list(range(self.n - 1, -1, -1))).reshape(-1)

    # -- bond read-outs (mps_core.py:465-494) -----------------------------------
    def entanglement_entropy(self, bond: int) -> float:
        """-sum p log p over the squared singular values (above 1e-15, normalised) of
        sites 0..bond contracted into a 2^(bond+1) x chi matrix (mps_core.py:465-483).
        On the device (MPSChain.entanglement_entropies) while this copy is unchanged
        from its device chain; otherwise from the host tensors."""
        if bond < 0 or bond >= self.n - 1:
            raise ValueError(f"bond {bond} out of range [0, {self.n - 2}]")
        if self.device_chain is not None:
            return self.device_chain.entanglement_entropy(bond)
        acc = self.sites[0][0]
        for t in self.sites[1:bond + 1]:
            acc = np.tensordot(acc, t, axes=([1], [0])).reshape(-1, t.shape[2])
        sv = np.linalg.svd(acc, compute_uv=False)
        sv = sv[sv > 1e-15]
        p = sv * sv
        p = p / p.sum()
        return -float(np.sum(p * np.log(np.maximum(p, 1e-300))))

    def bond_spectrum(self, bond: int) -> np.ndarray:
        """Singular values of site `bond` as a (2 chi_l) x chi_r matrix after sites
        0..bond of a copy are left-normalised, that site included, as the reference
        does (mps_core.py:485-494).  From the device bond dimensions while this copy is
        unchanged from its device chain (MPSChain.bond_spectrum)."""
        if self.device_chain is not None:
            return self.device_chain.bond_spectrum(bond)
        m = self.copy()
        for i in range(bond + 1):
            m.left_normalize_site(i)
        a = m.sites[bond]
        return np.linalg.svd(a.reshape(-1, a.shape[2]), compute_uv=False)

    # -- shots (mps_core.py:371-431), drawn on the device ------------------------
    def sample(self, num_shots: int = 1, measured_qubits: Optional[List[int]] = None,
               rng: Optional[np.random.Generator] = None, device: int = 0) -> List[str]:
        return MPSChain.from_host(self, device).sample(num_shots, measured_qubits, rng)

    def get_counts(self, num_shots: int, measured_qubits: Optional[List[int]] = None,
                   rng: Optional[np.random.Generator] = None, device: int = 0) -> Dict[str, int]:
        return MPSChain.from_host(self, device).get_counts(num_shots, measured_qubits, rng)


def z_expectations(host: 'HostMPS') -> np.ndarray:
    """<Z_i> / <psi|psi> for every site of a host chain, by left and right transfer
    matrices (normalised at every step; O(n chi^3))."""
    n = host.n
    right = [None] * (n + 1)
    right[n] = np.ones((1, 1), dtype=np.complex128)
    for i in range(n - 1, -1, -1):
        a = host.sites[i]
        t = np.tensordot(a, right[i + 1], axes=([2], [0]))          # (l, s, r')
        r = np.tensordot(t, a.conj(), axes=([1, 2], [1, 2]))        # (l, l')
        right[i] = r / max(abs(np.trace(r)), 1e-300)
    out = np.zeros(n)
    left = np.ones((1, 1), dtype=np.complex128)
    z = np.array([1.0, -1.0])
    for i in range(n):
        a = host.sites[i]
        t = np.tensordot(left, a, axes=([1], [0]))                  # (l', s, r)
        u = np.tensordot(t, right[i + 1], axes=([2], [0]))          # (l', s, r')
        w = np.tensordot(a.conj(), u, axes=([0, 2], [0, 2]))        # (s*, s)
        num = float(np.real(z[0] * w[0, 0] + z[1] * w[1, 1]))
        den = float(np.real(w[0, 0] + w[1, 1]))
        out[i] = num / den if den != 0 else 0.0
        nl = np.tensordot(a.conj(), t, axes=([0, 1], [0, 1]))       # (r*, r)
        left = nl / max(abs(np.trace(nl)), 1e-300)
    return out


def overlap(bra, ket) -> complex:
    """<bra|ket> of two chains (host copies) by transfer matrices, with exp(norm_log)."""
    env = np.ones((1, 1), dtype=np.complex128)
    for a, b in zip(bra.sites, ket.sites):
        # env[x, y] a*[x, s, p] b[y, s, q] -> new env[p, q]
        t = np.tensordot(env, b, axes=([1], [0]))              # (x, s, q)
        env = np.tensordot(a.conj(), t, axes=([0, 1], [0, 1]))  # (p, q)
    return complex(env[0, 0]) * math.exp(bra.norm_log + ket.norm_log)
