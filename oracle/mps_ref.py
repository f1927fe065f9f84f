"""CPU oracle for the MPS gate-application path (test infrastructure only).

This module is the checker for the HIP MPS path: it is imported by `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`, and by
nothing in the product package.  It restates, in NumPy/SciPy FP64, the
semantics of the reference's pure-Python MPS (`mps_core.py`) and of the
MPS host driver's circuit compiler and gate loop (`fpga_mps_simulator.py`).

Pinning (what checks each function; see tests/test_oracle_mps.py):
  * gate_1q / gate_2q           closed forms + unitarity + textbook identities
  * swap_route                   exact statevectors of the routed KAT circuits
  * MPS.apply_2q / svd_truncate  brute-force contraction at tiny chi, LAPACK
                                 singular values, U^H U = I, keep rule, the
                                 norm_log closed form log(1-eps), and the
                                 golden vectors written from mps_core.py
                                 (tests/golden/mps_core_*.npz)
  * get_statevector / pauli_expectation_sv   SV oracle round trip, closed forms
  * compile_circuit               routed-circuit statevectors vs SV oracle
  * right_environments / sample   exact marginals (chi^2 vs |psi|^2)
  * builder-added (z_expectations, overlap, tfim_energy): cross-checked against
    the statevector path at n <= 16.
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.linalg as sla

# ----------------------------------------------------------------------------
# Gate matrices (mps_core.py:21-144).  Basis for 2q gates: |q_first q_second>,
# the first-listed qubit is the most significant bit of the row index.
# ----------------------------------------------------------------------------

_R2 = 1.0 / math.sqrt(2.0)

FIXED_GATES: Dict[str, np.ndarray] = {}


def _c(rows) -> np.ndarray:
    return np.array(rows, dtype=np.complex128)


FIXED_GATES.update({
    'id': np.eye(2, dtype=np.complex128),
    'h': _c([[_R2, _R2], [_R2, -_R2]]),
    'x': _c([[0, 1], [1, 0]]),
    'y': _c([[0, -1j], [1j, 0]]),
    'z': _c([[1, 0], [0, -1]]),
    's': _c([[1, 0], [0, 1j]]),
    'sdg': _c([[1, 0], [0, -1j]]),
    't': _c([[1, 0], [0, np.exp(0.25j * math.pi)]]),
    'tdg': _c([[1, 0], [0, np.exp(-0.25j * math.pi)]]),
    'sx': _c([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]),
    'sxdg': _c([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]]),
    'cx': _c([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
    'cz': np.diag([1, 1, 1, -1]).astype(np.complex128),
    'cy': _c([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]),
    'swap': _c([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
    'iswap': _c([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
    'dcx': _c([[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]]),
})

ONE_QUBIT_NAMES = frozenset({'id', 'h', 'x', 'y', 'z', 's', 'sdg', 't', 'tdg', 'sx',
                             'sxdg', 'rx', 'ry', 'rz', 'p', 'phase', 'u1', 'u2',
                             'u3', 'u'})                         # mps_core.py:501-504
TWO_QUBIT_NAMES = frozenset({'cx', 'cnot', 'cy', 'cz', 'ch', 'cp', 'cphase', 'crx',
                             'cry', 'crz', 'csx', 'swap', 'iswap', 'dcx', 'ecr',
                             'rxx', 'ryy', 'rzz'})               # mps_core.py:506-510


def gate_1q(name: str, params: Optional[Sequence[float]] = None) -> np.ndarray:
    """2x2 matrix of a named one-qubit gate (mps_core.py:46-77)."""
    key = name.lower()
    if key in FIXED_GATES and FIXED_GATES[key].shape == (2, 2):
        return FIXED_GATES[key]
    p = [float(v) for v in (params or [])]
    if key == 'rx':
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        return _c([[c, -1j * s], [-1j * s, c]])
    if key == 'ry':
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        return _c([[c, -s], [s, c]])
    if key == 'rz':
        ph = np.exp(0.5j * p[0])
        return _c([[np.conj(ph), 0], [0, ph]])
    if key in ('p', 'phase', 'u1'):
        return _c([[1, 0], [0, np.exp(1j * p[0])]])
    if key == 'u2':
        return _R2 * _c([[1, -np.exp(1j * p[1])],
                         [np.exp(1j * p[0]), np.exp(1j * (p[0] + p[1]))]])
    if key in ('u3', 'u'):
        th, ph, lam = p[0], p[1], p[2]
        c, s = math.cos(th / 2), math.sin(th / 2)
        return _c([[c, -np.exp(1j * lam) * s],
                   [np.exp(1j * ph) * s, np.exp(1j * (ph + lam)) * c]])
    raise ValueError(f"Unknown 1q gate: {name}")


def _controlled(u: np.ndarray) -> np.ndarray:
    m = np.eye(4, dtype=np.complex128)
    m[2:, 2:] = u
    return m


def gate_2q(name: str, params: Optional[Sequence[float]] = None) -> np.ndarray:
    """4x4 matrix of a named two-qubit gate (mps_core.py:80-144).

    'cnot' is listed as a two-qubit name by the reference (mps_core.py:507) but
    has no matrix there; the reference's _gate_2q raises for it, and so does
    this function, so callers see the same error.
    """
    key = name.lower()
    if key in FIXED_GATES and FIXED_GATES[key].shape == (4, 4):
        return FIXED_GATES[key]
    p = [float(v) for v in (params or [])]
    if key in ('cp', 'cphase'):
        return np.diag([1, 1, 1, np.exp(1j * p[0])]).astype(np.complex128)
    if key == 'ch':
        return _controlled(FIXED_GATES['h'])
    if key == 'crx':
        return _controlled(gate_1q('rx', p))
    if key == 'cry':
        return _controlled(gate_1q('ry', p))
    if key == 'crz':
        return _controlled(gate_1q('rz', p))
    if key == 'csx':
        return _controlled(FIXED_GATES['sx'])
    if key == 'ecr':
        return _R2 * _c([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]])
    if key in ('rxx', 'ryy'):
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        corner = -1j * s if key == 'rxx' else 1j * s
        return _c([[c, 0, 0, corner], [0, c, -1j * s, 0],
                   [0, -1j * s, c, 0], [corner, 0, 0, c]])
    if key == 'rzz':
        e = np.exp(0.5j * p[0])
        return np.diag([np.conj(e), e, e, np.conj(e)]).astype(np.complex128)
    raise ValueError(f"Unknown 2q gate: {name}")


SWAP_PERM = np.array([0, 2, 1, 3])


def reverse_2q(g: np.ndarray) -> np.ndarray:
    """Re-express a 4x4 gate with its two qubits exchanged (mps_core.py:319-322)."""
    return g[np.ix_(SWAP_PERM, SWAP_PERM)]


def swap_route(i: int, j: int, name, params):
    """Nearest-neighbour SWAP chain for a gate on (i, j) (mps_core.py:151-165).

    The far qubit walks down to i+1, the gate acts on (i, i+1), and the chain
    walks back up.  Returns [(name, [a, b], params), ...].
    """
    lo, hi = (i, j) if i < j else (j, i)
    seq = [('swap', [k - 1, k], None) for k in range(hi, lo + 1, -1)]
    seq.append((name, [lo, lo + 1], params))
    seq.extend(('swap', [k, k + 1], None) for k in range(lo + 1, hi))
    return seq


# ----------------------------------------------------------------------------
# MPS container (mps_core.py:172-494)
# ----------------------------------------------------------------------------

class MPS:
    """List of site tensors (chi_l, 2, chi_r), C order, plus norm_log.

    Follows mps_core.MPSState: |0..0> start, sqrt(S) absorbed into both
    neighbours after every two-site update, no canonical centre.
    """

    def __init__(self, n: int, chi_max: int = 256, svd_cutoff: float = 1e-12):
        self.n = n
        self.chi_max = chi_max
        self.svd_cutoff = svd_cutoff
        self.norm_log = 0.0
        self.sites: List[np.ndarray] = []
        for _ in range(n):
            t = np.zeros((1, 2, 1), dtype=np.complex128)
            t[0, 0, 0] = 1.0
            self.sites.append(t)
        # Diagnostics recorded per two-site update (not part of mps_core).
        self.trace: List[Tuple[int, int, float]] = []

    def copy(self) -> 'MPS':
        m = MPS(self.n, self.chi_max, self.svd_cutoff)
        m.sites = [t.copy() for t in self.sites]
        m.norm_log = self.norm_log
        return m

    def bond_dims(self) -> List[int]:
        return [t.shape[2] for t in self.sites[:-1]]

    def max_bond(self) -> int:
        return 1 if self.n <= 1 else max(self.bond_dims())

    # -- one-site gate (mps_core.py:267-273) --------------------------------
    def apply_1q(self, site: int, u: np.ndarray) -> None:
        a = self.sites[site]
        self.sites[site] = np.einsum('ts,asb->atb', u, a)

    # -- truncated SVD (mps_core.py:277-300) --------------------------------
    def svd_truncate(self, m: np.ndarray):
        """SVD + truncation + norm_log bookkeeping of mps_core._svd_truncate.

        keep = clip(#{S > cutoff*S[0]}, 1, chi_max).  When singular values are
        dropped, norm_log gains 0.5*log(1 - sum(S_drop^2)/max(sum(S^2),1e-300));
        then the kept S is renormalised and norm_log gains log(||S_kept||).
        The truncation weight therefore enters norm_log twice (once explicitly,
        once through ||S_kept||) -- reproduced as is.
        """
        u, s, vh = sla.svd(m, full_matrices=False)
        thr = self.svd_cutoff * s[0] if s.size else 0.0
        keep = int(np.clip(int(np.count_nonzero(s > thr)), 1, self.chi_max))
        dlog = 0.0
        if keep < s.size:
            dropped = float(np.sum(s[keep:] ** 2))
            if dropped > 0:
                dlog += 0.5 * math.log(1.0 - dropped / max(float(np.sum(s ** 2)), 1e-300))
        u = u[:, :keep]
        vh = vh[:keep, :]
        st = s[:keep]
        nrm = float(np.linalg.norm(st))
        if nrm > 1e-15:
            dlog += math.log(nrm)
            st = st / nrm
        self.norm_log += dlog
        return u, np.sqrt(st), vh, s, dlog

    # -- two-site gate (mps_core.py:304-350) ----------------------------------
    def apply_2q(self, i: int, j: int, g: np.ndarray) -> None:
        if abs(i - j) != 1:
            for nm, qs, _ in swap_route(i, j, '__gate__', None):
                self.apply_2q(qs[0], qs[1], FIXED_GATES['swap'] if nm == 'swap' else g)
            return
        if i > j:
            i, j = j, i
            g = reverse_2q(g)
        self.apply_adjacent_2q(i, g)

    def theta(self, i: int, g: np.ndarray) -> np.ndarray:
        """Gate-applied two-site matrix M (2*chi_l, 2*chi_r) (mps_core.py:331-344)."""
        a, b = self.sites[i], self.sites[i + 1]
        cl, cr = a.shape[0], b.shape[2]
        two = np.tensordot(a, b, axes=([2], [0])).reshape(cl, 4, cr)
        two = np.einsum('st,atd->asd', g.reshape(4, 4), two)
        return two.reshape(2 * cl, 2 * cr)

    def apply_adjacent_2q(self, i: int, g: np.ndarray) -> None:
        a, b = self.sites[i], self.sites[i + 1]
        cl, cr = a.shape[0], b.shape[2]
        m = self.theta(i, g)
        u, sq, vh, s_full, dlog = self.svd_truncate(m)
        k = sq.size
        self.sites[i] = (u * sq[None, :]).reshape(cl, 2, k)
        self.sites[i + 1] = (sq[:, None] * vh).reshape(k, 2, cr)
        self.trace.append((i, k, dlog))

    # -- environments and sampling (mps_core.py:354-429) ---------------------
    def right_environments(self, stop: int = 0) -> List[np.ndarray]:
        """R_i for i >= stop (entries below `stop` stay None; 0 gives all of them)."""
        envs: List[Optional[np.ndarray]] = [None] * (self.n + 1)
        envs[self.n] = np.ones((1, 1), dtype=np.complex128)
        for i in range(self.n - 1, stop - 1, -1):
            a = self.sites[i]
            r = np.einsum('asb,bc,dsc->ad', a, envs[i + 1], a.conj())
            tr = float(np.real(np.trace(r)))
            if tr > 1e-300:
                r = r / tr
            envs[i] = r
        return envs  # type: ignore[return-value]

    def sample(self, shots: int, measured: Optional[List[int]] = None,
               rng: Optional[np.random.Generator] = None,
               envs: Optional[List[np.ndarray]] = None) -> List[str]:
        """Sequential conditional sampling (mps_core.py:371-420); bitstring[0] is qubit n-1.

        With a measured subset the key is built as mps_core.py:415-417 builds it: for q
        in the measured sites in descending order, the outcome of site n-1-q (the
        mirrored site, not site q itself).  `envs` replaces the right environments
        (right_environments() when None).
        """
        rng = rng if rng is not None else np.random.default_rng()
        measured = list(range(self.n)) if measured is None else list(measured)
        envs = self.right_environments() if envs is None else envs
        out = []
        full = set(measured) == set(range(self.n))
        for _ in range(shots):
            v = np.ones(1, dtype=np.complex128)
            bits = [0] * self.n
            for i in range(self.n):
                a = self.sites[i]
                r = envs[i + 1]
                pr = np.zeros(2)
                for sgm in (0, 1):
                    w = v @ a[:, sgm, :]
                    pr[sgm] = np.real(w @ r @ w.conj())
                pr = np.maximum(pr, 0.0)
                tot = pr.sum()
                pr = np.array([0.5, 0.5]) if tot < 1e-15 else pr / tot
                o = int(rng.choice(2, p=pr))
                bits[i] = o
                v = v @ a[:, o, :]
                nv = np.linalg.norm(v)
                if nv > 1e-15:
                    v = v / nv
            if full:
                out.append(''.join(str(bits[q]) for q in range(self.n - 1, -1, -1)))
            else:
                out.append(''.join(str(bits[self.n - 1 - q])
                                   for q in sorted(measured, reverse=True)))
        return out

    def get_counts(self, shots, measured=None, rng=None) -> Dict[str, int]:
        c: Dict[str, int] = {}
        for s in self.sample(shots, measured, rng):
            c[s] = c.get(s, 0) + 1
        return c

    # -- statevector (mps_core.py:433-461) -----------------------------------
    def get_statevector(self) -> np.ndarray:
        """Dense state, qubit q <-> index bit q (site 0 = least significant)."""
        if self.n > 24:
            raise ValueError(f"Statevector reconstruction for n={self.n} > 24 not supported.")
        acc = self.sites[0][0]                      # (2, chi)
        for i in range(1, self.n):
            acc = np.tensordot(acc, self.sites[i], axes=([1], [0]))
            acc = acc.reshape(-1, acc.shape[-1])
        sv = acc[:, 0].copy()                       # site 0 is the MSB here
        if self.norm_log != 0.0:
            sv *= math.exp(self.norm_log)
        return sv.reshape([2] * self.n).transpose(list(range(self.n - 1, -1, -1))).reshape(-1)

    # -- bond diagnostics (mps_core.py:465-494) ------------------------------
    def left_normalize_site(self, i: int) -> None:
        a = self.sites[i]
        cl, d, cr = a.shape
        q, r = np.linalg.qr(a.reshape(cl * d, cr))
        self.sites[i] = q.reshape(cl, d, q.shape[1])
        if i + 1 < self.n:
            self.sites[i + 1] = np.tensordot(r, self.sites[i + 1], axes=([1], [0]))

    def bond_spectrum(self, bond: int) -> np.ndarray:
        m = self.copy()
        for i in range(bond + 1):
            m.left_normalize_site(i)
        a = m.sites[bond]
        return sla.svd(a.reshape(-1, a.shape[2]), compute_uv=False)

    def entanglement_entropy(self, bond: int) -> float:
        if bond < 0 or bond >= self.n - 1:
            raise ValueError(f"bond {bond} out of range [0, {self.n - 2}]")
        acc = self.sites[0][0]
        for i in range(1, bond + 1):
            acc = np.tensordot(acc, self.sites[i], axes=([1], [0])).reshape(-1, self.sites[i].shape[2])
        s = sla.svd(acc, compute_uv=False)
        s = s[s > 1e-15]
        p = s ** 2 / np.sum(s ** 2)
        return -float(np.sum(p * np.log(np.maximum(p, 1e-300))))

    # -- builder-added observables (no mps_core counterpart) ------------------
    def norm_sq(self) -> float:
        """<psi|psi> including exp(2*norm_log), by transfer matrices."""
        return float(np.real(overlap(self, self)))

    def z_expectations(self) -> np.ndarray:
        """<Z_i>/<psi|psi> for every site, by left/right transfer matrices."""
        return local_expectations(self, FIXED_GATES['z'])


def _transfer_left(sites_a, sites_b):
    envs = [np.ones((1, 1), dtype=np.complex128)]
    for a, b in zip(sites_a, sites_b):
        envs.append(np.einsum('xy,xsa,ysb->ab', envs[-1], a.conj(), b))
    return envs


def overlap(bra: MPS, ket: MPS) -> complex:
    """<bra|ket> including both norm_log factors."""
    env = np.ones((1, 1), dtype=np.complex128)
    for a, b in zip(bra.sites, ket.sites):
        env = np.einsum('xy,xsa,ysb->ab', env, a.conj(), b)
    return complex(env[0, 0]) * math.exp(bra.norm_log + ket.norm_log)


def _environments(m: MPS):
    """Left and right transfer environments of <psi|psi>, each divided by its trace (the
    ratios below are unchanged, and long chains neither overflow nor underflow)."""
    n = m.n
    left = [np.ones((1, 1), dtype=np.complex128)]
    for a in m.sites:
        e = np.einsum('xy,xsa,ysb->ab', left[-1], a.conj(), a, optimize=True)
        left.append(e / max(abs(np.trace(e)), 1e-300))
    right = [None] * (n + 1)
    right[n] = np.ones((1, 1), dtype=np.complex128)
    for i in range(n - 1, -1, -1):
        a = m.sites[i]
        e = np.einsum('xsa,ysb,ab->xy', a.conj(), a, right[i + 1], optimize=True)
        right[i] = e / max(abs(np.trace(e)), 1e-300)
    return left, right


def local_expectations(m: MPS, op: np.ndarray) -> np.ndarray:
    """<O_i>/<psi|psi> for a one-site operator O on every site."""
    left, right = _environments(m)
    out = np.empty(m.n)
    eye = np.eye(2, dtype=np.complex128)
    for i in range(m.n):
        a = m.sites[i]
        v = np.einsum('xy,xsa,st,ytb,ab->', left[i], a.conj(), op, a, right[i + 1],
                      optimize=True)
        d = np.einsum('xy,xsa,st,ytb,ab->', left[i], a.conj(), eye, a, right[i + 1],
                      optimize=True)
        out[i] = np.real(v) / np.real(d)
    return out


def zz_neighbours(m: MPS) -> np.ndarray:
    """<Z_i Z_i+1>/<psi|psi> for i = 0..n-2."""
    left, right = _environments(m)
    z = FIXED_GATES['z']
    eye = np.eye(2, dtype=np.complex128)
    out = np.empty(m.n - 1)
    for i in range(m.n - 1):
        a, b = m.sites[i], m.sites[i + 1]
        num = np.einsum('xy,xsa,st,ytb->ab', left[i], a.conj(), z, a, optimize=True)
        num = np.einsum('xy,xsa,st,ytb->ab', num, b.conj(), z, b, optimize=True)
        den = np.einsum('xy,xsa,ysb->ab', left[i], a.conj(), a, optimize=True)
        den = np.einsum('xy,xsa,ysb->ab', den, b.conj(), b, optimize=True)
        out[i] = np.real(np.sum(num * right[i + 2])) / np.real(np.sum(den * right[i + 2]))
    return out


def random_product_state(n: int, seed: int) -> np.ndarray:
    """n normalised one-qubit states (n x 2 complex) from a seeded generator."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def product_overlap_log(m: MPS, phi: np.ndarray):
    """<phi|psi> for the product state phi (n x 2) as (log |<phi|psi>|, arg <phi|psi>),
    norm_log included; the running boundary vector is renormalised at every site."""
    v = np.ones(1, dtype=np.complex128)
    logabs = m.norm_log
    for i, a in enumerate(m.sites):
        v = np.einsum('s,x,xsb->b', phi[i].conj(), v, a)
        nv = float(np.linalg.norm(v))
        if nv == 0.0:
            return -math.inf, 0.0
        logabs += math.log(nv)
        v = v / nv
    return logabs + math.log(abs(v[0])), float(np.angle(v[0]))


def tfim_energy(m: MPS, j_coupling: float = 1.0, h_field: float = 1.0) -> float:
    """<H>/<psi|psi> for H = -J sum Z_i Z_{i+1} - h sum X_i (open chain)."""
    n = m.n
    z, x = FIXED_GATES['z'], FIXED_GATES['x']
    left = [np.ones((1, 1), dtype=np.complex128)]
    for a in m.sites:
        left.append(np.einsum('xy,xsa,ysb->ab', left[-1], a.conj(), a, optimize=True))
    right = [None] * (n + 1)
    right[n] = np.ones((1, 1), dtype=np.complex128)
    for i in range(n - 1, -1, -1):
        a = m.sites[i]
        right[i] = np.einsum('xsa,ysb,ab->xy', a.conj(), a, right[i + 1], optimize=True)
    nrm = np.real(left[n][0, 0])
    e = 0.0
    for i in range(n):
        a = m.sites[i]
        e -= h_field * np.real(np.einsum('xy,xsa,st,ytb,ab->', left[i], a.conj(), x, a, right[i + 1], optimize=True))
    for i in range(n - 1):
        a, b = m.sites[i], m.sites[i + 1]
        env = np.einsum('xy,xsa,st,ytb->ab', left[i], a.conj(), z, a, optimize=True)
        env = np.einsum('xy,xsa,st,ytb->ab', env, b.conj(), z, b, optimize=True)
        e -= j_coupling * np.real(np.einsum('ab,ab->', env, right[i + 2], optimize=True))
    return float(e / nrm)


def pauli_expectation_sv(sv: np.ndarray, pauli: str) -> float:
    """<psi|P|psi> for a Pauli string; character k acts on qubit n-1-k
    (fpga_mps_simulator.py:1091-1108)."""
    n = int(round(math.log2(sv.size)))
    psi = sv.reshape([2] * n)                 # axis 0 is qubit n-1
    out = psi.copy()
    for k, ch in enumerate(pauli.upper()):
        if ch == 'I':
            continue
        mat = FIXED_GATES.get(ch.lower(), np.eye(2, dtype=np.complex128))
        out = np.moveaxis(np.tensordot(mat, out, axes=([1], [k])), 0, k)
    return float(np.real(np.vdot(sv, out.reshape(-1))))


# ----------------------------------------------------------------------------
# Circuit compiler and gate loop (fpga_mps_simulator.py:272-410, 528-643)
# ----------------------------------------------------------------------------

MPS_OP_1Q = {'id': 0, 'h': 1, 'x': 2, 'y': 3, 'z': 4, 's': 5, 'sdg': 6, 't': 7, 'tdg': 8,
             'sx': 9, 'sxdg': 10, 'rx': 11, 'ry': 12, 'rz': 13, 'p': 14, 'phase': 14,
             'u1': 14, 'u2': 15, 'u3': 16, 'u': 16}          # fpga_mps_simulator.py:88-93
MPS_OP_2Q = {'cx': 17, 'cnot': 17, 'cy': 18, 'cz': 19, 'ch': 20, 'cp': 21, 'cphase': 21,
             'crx': 22, 'cry': 23, 'crz': 24, 'csx': 25, 'swap': 26, 'iswap': 27,
             'dcx': 28, 'ecr': 29, 'rxx': 30, 'ryy': 31, 'rzz': 32}


def ccx_sequence(c0: int, c1: int, t: int):
    """15-gate Clifford+T Toffoli (fpga_mps_simulator.py:272-282)."""
    return [('h', [t], None), ('cx', [c1, t], None), ('tdg', [t], None), ('cx', [c0, t], None),
            ('t', [t], None), ('cx', [c1, t], None), ('tdg', [t], None), ('cx', [c0, t], None),
            ('t', [c1], None), ('t', [t], None), ('h', [t], None), ('cx', [c0, c1], None),
            ('t', [c0], None), ('tdg', [c1], None), ('cx', [c0, c1], None)]


def circuit_ops(circuit) -> List[Tuple[str, List[int], Optional[List[float]]]]:
    """(name, qubit indices, params) from a duck-typed circuit object."""
    ops = []
    for inst in circuit.data:
        qs = []
        for q in inst.qubits:
            try:
                qs.append(circuit.find_bit(q).index)
            except Exception:
                qs.append(circuit.qubits.index(q))
        ps = [float(v) for v in getattr(inst.operation, 'params', [])] or None
        ops.append((inst.operation.name.lower(), qs, ps))
    return ops


def compile_ops(ops, n: int):
    """Nearest-neighbour gate list + measured sites (fpga_mps_simulator.py:285-410).

    ops: [(name, qubits, params)].  Returns (gates, measured) with gates as
    dicts {'type', 'name', 'sites', 'params', 'mat'}; physical sites follow the
    permutation that the SWAP chains leave behind.
    """
    perm = list(range(n))
    out: list = []
    measured: List[int] = []

    def emit_swap(a, b):
        lo, hi = min(a, b), max(a, b)
        out.append({'type': '2q', 'name': 'swap', 'sites': [lo, hi], 'params': None,
                    'mat': FIXED_GATES['swap']})
        for k in range(n):
            if perm[k] == lo:
                perm[k] = hi
            elif perm[k] == hi:
                perm[k] = lo

    def emit_2q(pa, pb, nm, ps):
        if abs(pa - pb) == 1:
            g = gate_2q(nm, ps)
            out.append({'type': '2q', 'name': nm, 'sites': [min(pa, pb), max(pa, pb)],
                        'params': ps, 'mat': reverse_2q(g) if pa > pb else g})
            return
        lo, hi = (pa, pb) if pa < pb else (pb, pa)
        g = reverse_2q(gate_2q(nm, ps)) if pa > pb else gate_2q(nm, ps)
        for k in range(hi, lo + 1, -1):
            emit_swap(k - 1, k)
        out.append({'type': '2q', 'name': nm, 'sites': [lo, lo + 1], 'params': ps, 'mat': g})
        for k in range(lo + 1, hi):
            emit_swap(k, k + 1)

    for nm, qs, ps in ops:
        if nm == 'measure':
            for q in qs:
                if perm[q] not in measured:
                    measured.append(perm[q])
            continue
        if nm in ('barrier', 'snapshot', 'delay', 'id'):
            continue
        if nm == 'reset':
            out.append({'type': 'reset', 'name': 'reset', 'sites': [perm[qs[0]]],
                        'params': None, 'mat': None})
            continue
        if nm in ONE_QUBIT_NAMES or nm in MPS_OP_1Q:
            out.append({'type': '1q', 'name': nm, 'sites': [perm[qs[0]]], 'params': ps,
                        'mat': gate_1q(nm, ps)})
            continue
        if nm in TWO_QUBIT_NAMES or nm in MPS_OP_2Q:
            emit_2q(perm[qs[0]], perm[qs[1]], nm, ps)
            continue
        if nm in ('ccx', 'toffoli', 'ccnot'):
            # The reference feeds already-permuted sites into the decomposition
            # and does not permute the sub-gates again (fpga_mps_simulator.py:383-404).
            # An adjacent sub-CX whose control site is above its target site is
            # emitted with the unreversed matrix there (:395-402), i.e. with
            # control and target exchanged; that behaviour is kept here.
            for snm, sqs, sps in ccx_sequence(perm[qs[0]], perm[qs[1]], perm[qs[2]]):
                if snm in ONE_QUBIT_NAMES:
                    out.append({'type': '1q', 'name': snm, 'sites': [sqs[0]], 'params': sps,
                                'mat': gate_1q(snm, sps)})
                else:
                    a, b = sqs
                    if abs(a - b) == 1:
                        out.append({'type': '2q', 'name': snm, 'sites': [min(a, b), max(a, b)],
                                    'params': sps, 'mat': gate_2q(snm, sps)})
                    else:
                        emit_2q(a, b, snm, sps)
            continue
    if not measured:
        measured = list(range(n))
    return out, sorted(measured)


def run_gate_list(n: int, gates, chi_max: int, cutoff: float) -> MPS:
    """Gate loop of _run_circuit_fpga in FP64 (fpga_mps_simulator.py:528-637)."""
    m = MPS(n, chi_max, cutoff)
    for g in gates:
        if g['type'] == 'reset':
            s = g['sites'][0]
            t = m.sites[s].copy()
            t[:, 1, :] = 0.0
            nrm = np.linalg.norm(t)
            if nrm > 1e-15:
                t /= nrm
            else:
                t = np.zeros_like(t)
                t[0, 0, 0] = 1.0
            m.sites[s] = t
        elif g['type'] == '1q':
            m.apply_1q(g['sites'][0], g['mat'])
        else:
            a, b = g['sites']
            assert b == a + 1
            m.apply_adjacent_2q(a, g['mat'])
    return m


def update_pair(args):
    """One two-site update of an adjacent pair given on its own: (A, B, gate, chi_max,
    cutoff) -> (A', B', norm_log increment), as MPS.apply_adjacent_2q does it in a chain.
    The work item of the process-parallel CPU baseline (the updates of one brick-wall
    layer are independent)."""
    a, b, g, chi_max, cutoff = args
    two = MPS(2, chi_max, cutoff)
    two.sites = [a, b]
    two.apply_adjacent_2q(0, g)
    return two.sites[0], two.sites[1], two.norm_log


# ----------------------------------------------------------------------------
# Workload generators (inputs only; shared by tests and bench)
# ----------------------------------------------------------------------------

def haar_4x4(rng: np.random.Generator) -> np.ndarray:
    """Haar-random 4x4 unitary: QR of a complex Gaussian, phase-fixed."""
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))[None, :]


def near_identity_4x4(rng: np.random.Generator, eps: float = 0.1) -> np.ndarray:
    """exp(-i*eps*H), H = (M + M^H)/2 with M complex Gaussian."""
    mm = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    h = 0.5 * (mm + mm.conj().T)
    return sla.expm(-1j * eps * h)


def brickwall_bonds(n: int, depth: int) -> List[List[int]]:
    """Bond lists per layer: even layers (0,1),(2,3).., odd layers (1,2),(3,4).."""
    return [list(range(layer % 2, n - 1, 2)) for layer in range(depth)]
