"""Many independent MPS circuits of uneven length and bond dimension, batched.

Each circuit is a brick-wall of two-site gates on its own chain (own n, chi_max,
depth, seed).  Round r submits layer r of every circuit whose depth exceeds r as
one qs_mps_update_batch call (split into chunks of at most `max_tasks` tasks so
the theta workspace stays bounded); within a chunk the library groups tasks by
SVD size class.  Task tables are built as NumPy structured arrays with the
qs_mps_task layout, and gates are generated per (circuit, layer) in vectorised
form, so the host keeps up with the device at millions of updates.

Graph mode (build_graph / replay) plans every round on the host first
(qs_mps_batch_plan), uploads all launch tables as one device buffer, and records
the launches of the whole run (qs_mps_batch_launch, which touches no host memory)
into one CUDA/HIP graph; a replay is then a single submission for all rounds.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from .mps import MPSChain, PRECISIONS

TASK_DTYPE = np.dtype([('sites', '<u8'), ('bond', '<u8'), ('dlog', '<u8'),
                       ('site_stride', '<i8'), ('site', '<i4'), ('chi_max', '<i4'),
                       ('cutoff', '<f8'), ('n_sites', '<i4'), ('reserved', '<i4'),
                       ('gate', '<f8', (32,))])
assert TASK_DTYPE.itemsize == ctypes.sizeof(_lib.QsMpsTask)


def near_identity_batch(rng: np.random.Generator, k: int, eps: float = 0.1) -> np.ndarray:
    """k gates exp(-i*eps*H), H = (M + M^H)/2, M complex Gaussian: (k, 4, 4)."""
    m = rng.standard_normal((k, 4, 4)) + 1j * rng.standard_normal((k, 4, 4))
    h = 0.5 * (m + np.conj(np.swapaxes(m, 1, 2)))
    w, v = np.linalg.eigh(h)
    return (v * np.exp(-1j * eps * w)[:, None, :]) @ np.conj(np.swapaxes(v, 1, 2))


@dataclass
class CircuitSpec:
    n: int
    chi_max: int
    depth: int
    seed: int


def random_specs(count: int = 256, seed: int = 7, n_range=(50, 2000),
                 chis=(16, 32, 64), depth_range=(10, 60)) -> List[CircuitSpec]:
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        out.append(CircuitSpec(int(rng.integers(n_range[0], n_range[1] + 1)),
                               int(rng.choice(chis)),
                               int(rng.integers(depth_range[0], depth_range[1] + 1)),
                               int(seed * 100003 + c)))
    return out


class MultiRun:
    """Runs a list of circuit specs on one GPU, round by round."""

    def __init__(self, specs: List[CircuitSpec], cutoff: float = 1e-10,
                 precision: str = 'fp32', device: int = 0, max_tasks: int = 16384):
        import torch
        self.specs = specs
        self.cutoff = cutoff
        self.prec = PRECISIONS[precision]
        self.device = device
        self.max_tasks = max_tasks
        self.chains = [MPSChain(s.n, s.chi_max, cutoff, device, precision) for s in specs]
        self.rngs = [np.random.default_rng(s.seed) for s in specs]
        total = sum((s.n // 2) * ((s.depth + 1) // 2) + ((s.n - 1) // 2) * (s.depth // 2)
                    for s in specs)
        self.total_updates = total
        dev = f'cuda:{device}'
        self.dlog = torch.zeros(total, dtype=torch.float64, device=dev)
        self.owner = np.empty(total, dtype=np.int32)      # circuit of each update
        self._next = 0
        chi = max(s.chi_max for s in specs)
        nbytes = _lib.lib().qs_mps_workspace_bytes(max_tasks, chi)
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.status = torch.zeros(max_tasks, dtype=torch.int32, device=dev)
        self.err = torch.zeros(1, dtype=torch.int32, device=dev)
        self.rounds = max(s.depth for s in specs)

    def _round_tasks(self, r: int) -> np.ndarray:
        parts = []
        dbase = self.dlog.data_ptr()
        for c, (s, ch) in enumerate(zip(self.specs, self.chains)):
            if r >= s.depth:
                continue
            sites = np.arange(r % 2, s.n - 1, 2, dtype=np.int32)
            k = sites.size
            if k == 0:
                continue
            g = near_identity_batch(self.rngs[c], k)
            t = np.zeros(k, dtype=TASK_DTYPE)
            t['sites'] = ch.sites.data_ptr()
            t['bond'] = ch.bond.data_ptr()
            t['dlog'] = dbase + 8 * (self._next + np.arange(k, dtype=np.uint64))
            t['site_stride'] = ch.stride
            t['site'] = sites
            t['n_sites'] = s.n
            t['chi_max'] = s.chi_max
            t['cutoff'] = self.cutoff
            gf = g.reshape(k, 16)
            t['gate'][:, 0::2] = gf.real
            t['gate'][:, 1::2] = gf.imag
            self.owner[self._next:self._next + k] = c
            self._next += k
            parts.append(t)
        return np.concatenate(parts) if parts else np.zeros(0, dtype=TASK_DTYPE)

    def _submit(self, tasks: np.ndarray) -> None:
        import torch
        stream = _lib.current_stream(self.device)
        for b in range(0, tasks.size, self.max_tasks):
            chunk = np.ascontiguousarray(tasks[b:b + self.max_tasks])
            ptr = chunk.ctypes.data_as(ctypes.POINTER(_lib.QsMpsTask))
            _lib.call('qs_mps_update_batch', ptr, chunk.size, self.prec,
                      ctypes.c_void_p(self.ws.data_ptr()), self.ws.numel(),
                      ctypes.c_void_p(self.status.data_ptr()), stream)
            self.err = torch.minimum(self.err, self.status[:chunk.size].min().reshape(1))

    def run(self, rounds: Optional[int] = None) -> None:
        for r in range(self.rounds if rounds is None else rounds):
            self._submit(self._round_tasks(r))

    # -- graph mode --------------------------------------------------------------
    def build_graph(self, rounds: Optional[int] = None, progress=None):
        """Plan every round, upload the launch tables and record the run as graphs, one
        per round of updates.

        The graphs cover exactly what run() would do for the same rounds (same gates,
        same dlog slots); replay() executes them in order.  `progress(r, rounds)` is
        called after each round is planned.  Returns the list of torch CUDAGraphs.
        """
        import torch
        lib = _lib.lib()
        chunks = []                  # (table byte offset, ntasks, counts, chi, status offset)
        tables = []
        round_of_chunk = []
        off = 0
        soff = 0
        nround = self.rounds if rounds is None else rounds
        for r in range(nround):
            tasks = self._round_tasks(r)
            for b in range(0, tasks.size, self.max_tasks):
                chunk = np.ascontiguousarray(tasks[b:b + self.max_tasks])
                nb = lib.qs_mps_batch_table_bytes(chunk.size)
                buf = np.zeros(nb, dtype=np.uint8)
                counts = (ctypes.c_int32 * 5)()
                chi = ctypes.c_int(0)
                _lib.call('qs_mps_batch_plan', ctypes.c_void_p(chunk.ctypes.data), chunk.size,
                          ctypes.c_void_p(buf.ctypes.data), nb, counts, ctypes.byref(chi))
                tables.append(buf)
                chunks.append((off, chunk.size, counts, chi.value, soff))
                round_of_chunk.append(r)
                off += nb
                soff += chunk.size
            if progress is not None:
                progress(r, nround)
        dev = f'cuda:{self.device}'
        host = np.concatenate(tables) if tables else np.zeros(1, dtype=np.uint8)
        self.g_tables = torch.from_numpy(host).to(dev)
        self.g_status = torch.zeros(max(soff, 1), dtype=torch.int32, device=dev)
        self.g_chunks = chunks
        torch.cuda.synchronize(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        tbase = self.g_tables.data_ptr()
        sbase = self.g_status.data_ptr()
        graphs = []
        for r in sorted(set(round_of_chunk)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=side):
                stream = _lib.current_stream(self.device)
                for (toff, n, counts, chi, so), rc in zip(chunks, round_of_chunk):
                    if rc != r:
                        continue
                    _lib.call('qs_mps_batch_launch', ctypes.c_void_p(tbase + toff), n, counts,
                              chi, self.prec, ctypes.c_void_p(self.ws.data_ptr()),
                              self.ws.numel(), ctypes.c_void_p(sbase + 4 * so), stream)
            graphs.append(g)
        torch.cuda.current_stream(self.device).wait_stream(side)
        self.graphs = graphs
        return graphs

    def replay(self, progress=None) -> None:
        """Replay the recorded rounds in order; `progress(k, count)` after each launch."""
        for k, g in enumerate(self.graphs):
            g.replay()
            if progress is not None:
                progress(k, len(self.graphs))

    def check_graph(self) -> None:
        worst = int(self.g_status.min().item())
        if worst < 0:
            raise RuntimeError(f'a two-site update failed with status {worst}')

    def check(self) -> None:
        worst = int(self.err.cpu().item())
        if worst < 0:
            raise RuntimeError(f'a two-site update failed with status {worst}')

    def norm_logs(self) -> np.ndarray:
        """norm_log of every circuit (increments summed in update order)."""
        d = self.dlog.cpu().numpy()[:self._next]
        return np.bincount(self.owner[:self._next], weights=d, minlength=len(self.specs))
