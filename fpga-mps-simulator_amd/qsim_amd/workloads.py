"""Synthetic circuit inputs for benchmarks and tests (seeded, deterministic)."""

from __future__ import annotations

import math
from typing import List

import numpy as np


def qft_gates(n: int, inverse: bool = False) -> list:
    """QFT gate list, ordered as comprehensive_test_v07.py:433-452 builds it."""
    g = []
    if not inverse:
        for i in range(n):
            g.append(('h', [i], None))
            for j in range(i + 1, n):
                g.append(('cp', [j, i], [math.pi / (2 ** (j - i))]))
        for i in range(n // 2):
            g.append(('swap', [i, n - 1 - i], None))
    else:
        for i in range(n // 2):
            g.append(('swap', [i, n - 1 - i], None))
        for i in range(n - 1, -1, -1):
            for j in range(n - 1, i, -1):
                g.append(('cp', [j, i], [-math.pi / (2 ** (j - i))]))
            g.append(('h', [i], None))
    return g


def u3_cz_layers(n: int, depth: int, seed: int) -> list:
    """depth x [U3(theta, phi, lambda ~ U[0, 2pi)) on every qubit, CZ(i, i+1) ladder]."""
    rng = np.random.default_rng(seed)
    g = []
    for _ in range(depth):
        ang = rng.uniform(0.0, 2 * math.pi, size=(n, 3))
        for q in range(n):
            g.append(('u3', [q], [float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2])]))
        for q in range(n - 1):
            g.append(('cz', [q, q + 1], None))
    return g


def sv_benchmark_circuit(n: int = 28, depth: int = 20, seed: int = 1234) -> list:
    """The 28-qubit statevector workload: U3 + CZ-ladder layers, then a QFT."""
    return u3_cz_layers(n, depth, seed) + qft_gates(n)


def inverse_circuit(gates: list) -> list:
    """Gate list of the inverse circuit (for round-trip checks)."""
    inv = {'s': 'sdg', 'sdg': 's', 't': 'tdg', 'tdg': 't', 'sx': 'sxdg', 'sxdg': 'sx'}
    out = []
    for name, qs, ps in reversed(gates):
        nm = name.lower()
        if nm in inv:
            out.append((inv[nm], qs, ps))
        elif nm in ('rx', 'ry', 'rz', 'p', 'phase', 'u1', 'cp', 'cphase', 'crx', 'cry', 'crz',
                    'rxx', 'ryy', 'rzz'):
            out.append((nm, qs, [-float(ps[0])]))
        elif nm in ('u3', 'u'):
            th, phi, lam = (float(v) for v in ps)
            out.append(('u3', qs, [-th, -lam, -phi]))
        elif nm == 'u2':
            phi, lam = (float(v) for v in ps)
            out.append(('u3', qs, [-math.pi / 2, -lam, -phi]))
        elif nm in ('h', 'x', 'y', 'z', 'cx', 'cnot', 'cy', 'cz', 'ch', 'swap', 'ccx', 'mcx',
                    'id'):
            out.append((nm, qs, ps))
        else:
            raise ValueError(f"no inverse rule for {name}")
    return out


# ------------------------------------------------------------------ MPS inputs

def haar_4x4(rng: np.random.Generator) -> np.ndarray:
    """Haar-random 4x4 unitary: QR of a complex Gaussian matrix, phase-fixed."""
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))[None, :]


def near_identity_4x4(rng: np.random.Generator, eps: float = 0.1) -> np.ndarray:
    """exp(-i*eps*H) with H = (M + M^H)/2, M complex Gaussian (eigen-decomposition)."""
    m = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    h = 0.5 * (m + m.conj().T)
    w, v = np.linalg.eigh(h)
    return (v * np.exp(-1j * eps * w)[None, :]) @ v.conj().T


def brickwall_bonds(n: int, depth: int) -> List[List[int]]:
    """Left sites of the bonds of each layer: (0,1),(2,3).. then (1,2),(3,4).."""
    return [list(range(layer % 2, n - 1, 2)) for layer in range(depth)]


def brickwall_gates(n: int, depth: int, seed: int, kind: str = 'near_identity') -> list:
    """[(layer, site, 4x4 gate)] for a brick-wall circuit."""
    rng = np.random.default_rng(seed)
    out = []
    for layer, bonds in enumerate(brickwall_bonds(n, depth)):
        for i in bonds:
            g = haar_4x4(rng) if kind == 'haar' else near_identity_4x4(rng)
            out.append((layer, i, g))
    return out
