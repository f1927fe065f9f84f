"""Autotuner for the two-site update's SVD tile shape.

The register SVD of size class 8, 16, 32 or 64 runs one team of `lanes` lanes per
column pair of theta; the library builds several widths per class
(qs_mps_svd_lane_choices).  For each class the tuner times every width on a batch
of full-size updates (a seeded chain whose interior bonds all have the class's
dimension, random site tensors, Haar two-site gates on every other bond; the chain
is restored before each timed run), and installs the fastest with
qs_mps_set_svd_lanes.  Results are kept per process and can be saved to / loaded
from a JSON file keyed by the device name, so a deployment tunes once.

Every width computes the same decomposition (parity tests run each one against the
oracle); only the summation order of the column inner products differs.
"""

from __future__ import annotations

import ctypes
import json
import os
from typing import Dict, Iterable, Optional

import numpy as np

from . import _lib
from .mps import MPSChain

CLASSES = (8, 16, 32, 64)

_tuned: Dict[int, int] = {}          # class -> lanes installed by this process


def lane_choices(chi_class: int) -> list:
    buf = (ctypes.c_int * 8)()
    cnt = ctypes.c_int(0)
    _lib.call('qs_mps_svd_lane_choices', chi_class, buf, 8, ctypes.byref(cnt))
    return [buf[i] for i in range(cnt.value)]


def current_lanes(chi_class: int) -> int:
    v = _lib.lib().qs_mps_svd_lanes(chi_class)
    if v < 0:
        _lib.check('qs_mps_svd_lanes', v)
    return v


def set_lanes(chi_class: int, lanes: int) -> None:
    _lib.call('qs_mps_set_svd_lanes', chi_class, lanes)


def _haar(rng: np.random.Generator, k: int) -> np.ndarray:
    z = (rng.standard_normal((k, 4, 4)) + 1j * rng.standard_normal((k, 4, 4))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    return q * (d / np.abs(d))[:, None, :]


def _full_bond_chain(chi: int, ntasks: int, precision: str, device: int, seed: int):
    """Chain of 2 ntasks + 2 sites, every interior bond = chi, random sites."""
    import torch
    n = 2 * ntasks + 2
    ch = MPSChain(n, chi, 1e-12, device, precision)
    b = np.full(n + 1, chi, dtype=np.int32)
    b[0] = b[n] = 1
    ch.bond.copy_(torch.from_numpy(b))
    gen = torch.Generator(device=ch.sites.device).manual_seed(seed)
    rows = ch.sites.view(n, ch.stride)
    vals = torch.randn(n, ch.stride, dtype=torch.complex128, device=ch.sites.device,
                       generator=gen)
    rows.copy_(vals)                   # slots hold (b_i, 2, b_i+1) in their first entries
    return ch


def _time_class(chi: int, lanes_list: Iterable[int], ntasks: int, reps: int,
                precision: str, device: int, seed: int) -> Dict[int, float]:
    import torch
    ch = _full_bond_chain(chi, ntasks, precision, device, seed)
    snap_sites, snap_bond = ch.sites.clone(), ch.bond.clone()
    gates = _haar(np.random.default_rng(seed), ntasks)
    updates = [(1 + 2 * j, gates[j]) for j in range(ntasks)]
    stream = torch.cuda.current_stream(device)
    out = {}
    for lanes in lanes_list:
        set_lanes(chi, lanes)
        best = float('inf')
        for r in range(reps + 1):          # first run warms up
            ch.sites.copy_(snap_sites)
            ch.bond.copy_(snap_bond)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ch.apply_2q_batch(updates)
            e1.record(stream)
            e1.synchronize()
            if r > 0:
                best = min(best, e0.elapsed_time(e1))
        ch.check_status()
        out[lanes] = best / ntasks
    return out


def autotune(classes: Iterable[int] = CLASSES, precision: str = 'fp32', device: int = 0,
             ntasks: Optional[Dict[int, int]] = None, reps: int = 3, seed: int = 11,
             verbose: bool = False) -> Dict[int, dict]:
    """Time every built team width per class; install the fastest.

    Returns {class: {'lanes': best, 'ms_per_update': {lanes: ms}}}.
    """
    sizes = {8: 4096, 16: 4096, 32: 2048, 64: 512}
    if ntasks:
        sizes.update(ntasks)
    report = {}
    for c in classes:
        times = _time_class(c, lane_choices(c), sizes[c], reps, precision, device, seed)
        best = min(times, key=times.get)
        set_lanes(c, best)
        _tuned[c] = best
        report[c] = {'lanes': best, 'ms_per_update': times}
        if verbose:
            print(f'chi class {c}: ' + ', '.join(f'{k} lanes {v * 1e3:.1f} us'
                                                for k, v in times.items()) +
                  f' -> {best}', flush=True)
    return report


def tuned() -> Dict[int, int]:
    return dict(_tuned)


def _device_key(device: int) -> str:
    import torch
    return torch.cuda.get_device_name(device)


def save(path: str, device: int = 0) -> None:
    """Write the widths in use to `path` under this device's name."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[_device_key(device)] = {str(c): current_lanes(c) for c in CLASSES}
    with open(path, 'w') as f:
        json.dump(data, f, indent=1, sort_keys=True)


def load(path: str, device: int = 0) -> bool:
    """Install widths saved for this device; False if the file has none."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        data = json.load(f)
    entry = data.get(_device_key(device))
    if not entry:
        return False
    for c, lanes in entry.items():
        set_lanes(int(c), int(lanes))
        _tuned[int(c)] = int(lanes)
    return True
