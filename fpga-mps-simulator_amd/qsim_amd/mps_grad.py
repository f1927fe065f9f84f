"""Gradients of an MPS energy through the two-site updates (reverse sweep over a tape).

Forward: a ParamCircuit (qsim_amd.adjoint) runs on the GPU MPS engine (MPSChain with the
FP64 contraction and Jacobi SVD kernels); every op records what its backward needs:

  * one-qubit rotation on site i: A' = M(theta) A            (records A and theta)
  * two-site gate on (i, i+1):    theta = G . (A B) as a (2 chi_l x 2 chi_r) matrix,
                                  theta = U S V^H truncated to k,
                                  A' = U sqrt(S_t), B' = sqrt(S_t) V^H, S_t = S / ||S_k||
                                  (records A, B, G and the outputs A', B')

Energy: E = <psi|H|psi> / <psi|psi> for the transverse-field Ising chain
H = -J sum Z_i Z_{i+1} - h sum X_i (open boundary), by left environments; torch autograd
of that contraction gives the cotangents of the final site tensors.

Backward (hand-written, torch tensor ops on the device).  Cotangent convention:
dL = Re sum conj(g) dz, i.e. g = dL/dRe z + i dL/dIm z (torch's convention).

  * rotation:  g_A = M^H g_A'   and   dE/dtheta = Re <g_A', (dM/dtheta) A>
  * two-site:  U, S_t, V are read back from A' and B' (the column norms of A' are
               sqrt(S_t)); S from U^H theta V;  g_U = g_A' D, g_V = g_B'^H D and
               g_S = D/(2S) [Re diag(U^H g_A') + Re diag(g_B' V)]  (D = diag sqrt(S_t));
               svd_backward (below) gives g_theta; the renormalisation S -> S_t only
               rescales the state and E is scale invariant, so it carries no gradient;
               contraction: g_T = G^H g_theta, g_A = g_T B^H, g_B = A^H g_T.

No counterpart in the reference (mps_core.py has no differentiation); the forward follows
mps_core.py:277-350 through the GPU kernels.
"""

from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import gates as _gates
from .adjoint import ParamCircuit

_PAULI_MAT = {
    'rx': np.array([[0, 1], [1, 0]], dtype=np.complex128),
    'ry': np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    'rz': np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def rotation(name: str, theta: float) -> np.ndarray:
    return _gates.matrix_1q(name, [float(theta)])


def drotation(name: str, theta: float) -> np.ndarray:
    """d/dtheta exp(-i theta P/2) = -(i/2) P exp(-i theta P/2)."""
    return -0.5j * _PAULI_MAT[name] @ rotation(name, theta)


# ---------------------------------------------------------------- SVD backward
def svd_backward(U, S, V, gU, gS, gV, eps: float = 1e-24):
    """Cotangent of theta for theta = U diag(S) V^H (thin or truncated, S > 0).

    U (m x k), S (k, real), V (n x k); gU (m x k), gV (n x k) complex cotangents, gS (k)
    real, with dL = Re<gU, dU> + gS . dS + Re<gV, dV>.  Returns g (m x n) with
    dL = Re<g, dtheta>.  From dP = U^H dtheta V:
        dS = Re diag dP,
        U^H dU = F o (dP S + S dP^H) + i diag(Im dP_kk / s_k),   V^H dV = F o (S dP + dP^H S)
    (the singular-vector phases are fixed by diag(V^H dV) = 0), plus the components of
    dU, dV outside span(U), span(V).  F_ij = 1/(s_j^2 - s_i^2) is broadened to
    (s_j^2 - s_i^2)/((s_j^2 - s_i^2)^2 + eps s_max^4) so that near-degenerate singular
    values give bounded terms (their contributions cancel for gauge-invariant losses).
    The default eps = 1e-24 broadens only gaps in s^2 below ~1e-12 s_max^2; the relative
    error it adds to a term elsewhere is eps s_max^4 / gap^2 (states of the 200-qubit
    ansatz have singular-value pairs with gaps near 1e-8, where eps = 1e-16 already
    shifts gradients by 1e-6).
    """
    import torch
    dt = U.dtype
    Uh = U.conj().transpose(-2, -1)
    Vh = V.conj().transpose(-2, -1)
    s2 = S * S
    diff = s2.unsqueeze(-2) - s2.unsqueeze(-1)          # [i, j] = s_j^2 - s_i^2
    smax4 = (s2.max() ** 2) if S.numel() else torch.ones((), dtype=S.dtype, device=S.device)
    F = diff / (diff * diff + eps * smax4)
    F.diagonal(dim1=-2, dim2=-1).zero_()
    Sinv = torch.where(S > 0, 1.0 / S, torch.zeros_like(S))
    Sc, Sic = S.to(dt), Sinv.to(dt)
    A = Uh @ gU
    B = Vh @ gV
    Fc = F.to(dt)
    X = (Fc * (A - A.conj().transpose(-2, -1))) * Sc.unsqueeze(-2)      # (F o (A - A^H)) S
    X = X + Sc.unsqueeze(-1) * (Fc * (B - B.conj().transpose(-2, -1)))  # S (F o (B - B^H))
    dA = torch.diagonal(A, dim1=-2, dim2=-1)
    X = X + torch.diag_embed(gS.to(dt) + 1j * (dA.imag * Sinv).to(dt))
    g = U @ X @ Vh
    m, n = U.shape[-2], V.shape[-2]
    Pu = torch.eye(m, dtype=dt, device=U.device) - U @ Uh
    Pv = torch.eye(n, dtype=dt, device=U.device) - V @ Vh
    g = g + Pu @ (gU * Sic.unsqueeze(-2)) @ Vh
    g = g + U @ (Sic.unsqueeze(-1) * gV.conj().transpose(-2, -1)) @ Pv
    return g


def two_site_backward(A, B, G, A2, B2, gA2, gB2, eps: float = 1e-24):
    """Cotangents of A (cl,2,cm) and B (cm,2,cr) for (A2, B2) = update(A, B, G).

    A2 is (cl,2,k), B2 (k,2,cr); G the 4x4 gate (row 2 s0 + s1, column 2 p + q).
    """
    import torch
    cl, cm = A.shape[0], A.shape[2]
    cr = B.shape[2]
    k = A2.shape[2]
    m, n = 2 * cl, 2 * cr
    # theta = G . (A B), rows (a, s0), columns (s1, d)
    T = torch.einsum('apb,bqd->apqd', A, B).reshape(cl, 4, cr)
    th = torch.einsum('st,atd->asd', G, T).reshape(cl, 2, 2, cr).reshape(m, n)
    Am = A2.reshape(m, k)
    Bm = B2.reshape(k, n)
    st = (Am.abs() ** 2).sum(dim=0)                       # S_t = column norms^2 of A'
    D = torch.sqrt(st)
    U = Am / D.to(Am.dtype).unsqueeze(0)
    Vh = Bm / D.to(Bm.dtype).unsqueeze(1)
    V = Vh.conj().transpose(0, 1)
    S = torch.real(torch.diagonal(U.conj().transpose(0, 1) @ th @ V))
    gAm = gA2.reshape(m, k)
    gBm = gB2.reshape(k, n)
    Dc = D.to(Am.dtype)
    gU = gAm * Dc.unsqueeze(0)
    gV = gBm.conj().transpose(0, 1) * Dc.unsqueeze(0)
    gS = D / (2.0 * S) * (torch.real(torch.diagonal(U.conj().transpose(0, 1) @ gAm)) +
                          torch.real(torch.diagonal(gBm @ V)))
    gth = svd_backward(U, S, V, gU, gS, gV, eps)
    gT = torch.einsum('st,asd->atd', G.conj(), gth.reshape(cl, 4, cr)).reshape(cl, 2, 2, cr)
    gA = torch.einsum('apqd,bqd->apb', gT, B.conj())
    gB = torch.einsum('apb,apqd->bqd', A.conj(), gT)
    return gA, gB


# ------------------------------------------------------------------ TFIM energy
def tfim_energy(sites: Sequence, J: float = 1.0, h: float = 1.0):
    """<psi|H|psi>/<psi|psi>, H = -J sum Z_i Z_{i+1} - h sum X_i, by left environments.

    Differentiable (torch); sites are (chi_l, 2, chi_r) complex tensors, site 0 = qubit 0.
    """
    import torch
    dt = sites[0].dtype
    dev = sites[0].device
    X = torch.tensor([[0, 1], [1, 0]], dtype=dt, device=dev)
    Z = torch.tensor([[1, 0], [0, -1]], dtype=dt, device=dev)
    one = torch.ones(1, 1, dtype=dt, device=dev)
    L, LZ, LH = one, None, torch.zeros(1, 1, dtype=dt, device=dev)

    def step(Lm, T, O=None):
        # new[b, b'] = sum conj(T[a, s, b]) Lm[a, a'] O[s, s'] T[a', s', b']
        if O is None:
            return torch.einsum('asb,ac,csd->bd', T.conj(), Lm, T)
        return torch.einsum('asb,ac,st,ctd->bd', T.conj(), Lm, O, T)

    for i, T in enumerate(sites):
        newLH = step(LH, T) - h * step(L, T, X)
        if LZ is not None:
            newLH = newLH - J * step(LZ, T, Z)
        newLZ = step(L, T, Z)
        L = step(L, T)
        LH, LZ = newLH, newLZ
    return torch.real(LH[0, 0] / L[0, 0])


# --------------------------------------------------------------- driver
class MPSEnergyGrad:
    """TFIM energy of a ParamCircuit run on the GPU MPS engine, and its gradient."""

    def __init__(self, circuit: ParamCircuit, chi_max: int = 32, cutoff: float = 1e-12,
                 J: float = 1.0, h: float = 1.0, device: int = 0, eps: float = 1e-24):
        self.c = circuit
        self.chi_max = chi_max
        self.cutoff = cutoff
        self.J = float(J)
        self.h = float(h)
        self.device = device
        self.eps = eps

    def _run(self, theta, record: bool):
        import torch
        from .mps import MPSChain
        ch = MPSChain(self.c.nq, self.chi_max, self.cutoff, self.device, precision='fp64')
        tape = []
        dev = ch.sites.device

        def site(i):
            b = ch.bond.cpu().numpy() if not record else bonds
            return ch.sites.view(ch.n, ch.stride)[i, :b[i] * 2 * b[i + 1]].reshape(
                int(b[i]), 2, int(b[i + 1]))

        bonds = np.ones(self.c.nq + 1, dtype=np.int64)
        for blk in self.c.blocks:
            if blk.kind == 'rot':
                # a rotation block acts on distinct qubits: one batched launch
                if record:
                    for name, q, k in blk.ops:
                        tape.append(('1q', q, name, k, site(q).clone()))
                ch.apply_1q_batch([(q, rotation(name, theta[k])) for name, q, k in blk.ops])
            else:
                batch: List[Tuple[int, np.ndarray]] = []
                busy = set()

                def flush():
                    nonlocal bonds
                    if not batch:
                        return
                    ins = [(i, site(i).clone(), site(i + 1).clone()) for i, _ in batch] \
                        if record else []
                    ch.apply_2q_batch(batch)
                    if record:
                        bonds = ch.bond.cpu().numpy().astype(np.int64)
                        for (i, g), (_, a0, b0) in zip(batch, ins):
                            tape.append(('2q', i, torch.as_tensor(g, device=dev), a0, b0,
                                         site(i).clone(), site(i + 1).clone()))
                    batch.clear()
                    busy.clear()

                for g in blk.ops:
                    name, qs = g[0].lower(), list(g[1])
                    if len(qs) != 2 or abs(qs[0] - qs[1]) != 1:
                        raise ValueError('fixed blocks of this driver hold adjacent 2q gates')
                    lo = min(qs)
                    mat = _gates.matrix_2q(name, g[2] if len(g) > 2 else None)
                    if qs[0] > qs[1]:
                        mat = _gates.swap_qubit_order(mat)
                    if lo in busy or lo + 1 in busy:
                        flush()
                    batch.append((lo, mat))
                    busy.update((lo, lo + 1))
                flush()
        ch.check_status()
        bonds = ch.bond.cpu().numpy().astype(np.int64)
        finals = [site(i).clone() for i in range(self.c.nq)]
        return finals, tape

    def energy(self, theta) -> float:
        import torch
        finals, _ = self._run(np.asarray(theta, dtype=np.float64), record=False)
        with torch.no_grad():
            return float(tfim_energy(finals, self.J, self.h))

    def energy_and_grad(self, theta) -> Tuple[float, np.ndarray]:
        theta = np.asarray(theta, dtype=np.float64)
        finals, tape = self._run(theta, record=True)
        return energy_and_grad_from_tape(finals, tape, theta, self.c.nparams, self.J,
                                         self.h, self.eps)


def torch_energy(module: MPSEnergyGrad):
    """A torch.autograd.Function computing E(theta) on the GPU MPS engine.

    The forward records the tape; the backward runs the reverse sweep through the
    recorded one-qubit rotations, two-site contractions and truncated SVDs.
    """
    import torch

    class _Energy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            th = theta.detach().cpu().numpy().astype(np.float64)
            finals, tape = module._run(th, record=True)
            ctx.finals, ctx.tape, ctx.theta = finals, tape, th
            with torch.no_grad():
                e = float(tfim_energy(finals, module.J, module.h))
            return torch.tensor(e, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            _, g = energy_and_grad_from_tape(ctx.finals, ctx.tape, ctx.theta,
                                             module.c.nparams, module.J, module.h,
                                             module.eps)
            ctx.finals = ctx.tape = None
            return grad_out * torch.as_tensor(g, dtype=grad_out.dtype, device=grad_out.device)

    return _Energy.apply


def backward_batch_device(records, cot, eps: float = 1e-24) -> None:
    """The reverse steps of independent two-site updates (disjoint bonds) in one launch of
    the hand-written kernel (qs_mps_backward_batch); updates cot in place.

    records: tape entries ('2q', i, G, A, B, A', B'); cot[i] holds the cotangent of site
    i after the update and receives the one before it."""
    import ctypes
    import torch
    from . import _lib
    k = len(records)
    dev = records[0][3].device
    chi = 1
    tasks = (_lib.QsMpsBwdTask * k)()
    outs = []
    keep = []
    gates = torch.stack([r[2] for r in records]).detach().cpu().numpy().reshape(k, 16)
    for j, (_, i, G, A, B, A2, B2) in enumerate(records):
        gA2 = cot[i].contiguous()
        gB2 = cot[i + 1].contiguous()
        gA = torch.empty_like(A)
        gB = torch.empty_like(B)
        keep += [gA2, gB2]
        outs.append((i, gA, gB))
        t = tasks[j]
        t.a, t.b, t.a2, t.b2 = A.data_ptr(), B.data_ptr(), A2.data_ptr(), B2.data_ptr()
        t.ga2, t.gb2, t.ga, t.gb = gA2.data_ptr(), gB2.data_ptr(), gA.data_ptr(), gB.data_ptr()
        t.cl, t.cm, t.cr, t.k = A.shape[0], A.shape[2], B.shape[2], A2.shape[2]
        t.eps = float(eps)
        t.gate[:] = [float(v) for z in gates[j] for v in (z.real, z.imag)]
        chi = max(chi, t.cl, t.cm, t.cr, t.k)
    nbytes = _lib.lib().qs_mps_backward_work_bytes(k, chi)
    work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    status = torch.full((k,), -99, dtype=torch.int32, device=dev)
    _lib.call('qs_mps_backward_batch', tasks, k, chi, ctypes.c_void_p(work.data_ptr()), nbytes,
              ctypes.c_void_p(status.data_ptr()), _lib.current_stream(dev.index))
    bad = int(status.min().item())
    if bad != 0:
        raise RuntimeError(f'two-site backward failed with status {bad}')
    for i, gA, gB in outs:
        cot[i], cot[i + 1] = gA, gB


def energy_and_grad_from_tape(finals, tape, theta, nparams: int, J: float = 1.0,
                              h: float = 1.0, eps: float = 1e-24,
                              device_kernels: Optional[bool] = None) -> Tuple[float, np.ndarray]:
    """E and dE/dtheta from the final site tensors and the forward tape (reverse sweep).

    Tape records: ('1q', site, rotation name, parameter index, A before) and
    ('2q', left site, G, A, B before, A', B' after).  With device_kernels (the default
    for a tape of GPU tensors, which is what MPSEnergyGrad records) the two-site steps run
    in the hand-written kernel, the records of disjoint bonds that follow each other in the
    tape (a brick-wall layer) in one launch; otherwise, and for tapes of host tensors,
    through two_site_backward (torch tensor ops, the same formulas).
    """
    import torch
    if device_kernels is None:
        device_kernels = bool(finals) and finals[0].is_cuda
    # Called from autograd.Function.backward too, where grad mode is off.
    with torch.enable_grad():
        leaves = [t.detach().requires_grad_(True) for t in finals]
        e = tfim_energy(leaves, J, h)
        gs = torch.autograd.grad(e, leaves)
    cot = {i: g for i, g in enumerate(gs)}
    grad = np.zeros(nparams)
    contrib: list = []          # (parameter index, device scalar) of the rotations
    pending: list = []
    busy: set = set()

    def flush():
        if pending:
            backward_batch_device(pending, cot, eps)
            pending.clear()
            busy.clear()

    for rec in reversed(tape):
        if rec[0] == '1q':
            _, q, name, k, A = rec
            if q in busy:
                flush()
            gA2 = cot[q]
            dM = torch.as_tensor(drotation(name, theta[k]), device=A.device)
            M = torch.as_tensor(rotation(name, theta[k]), device=A.device)
            dA2 = torch.einsum('st,atb->asb', dM, A)
            contrib.append((k, torch.real(torch.sum(gA2.conj() * dA2))))
            cot[q] = torch.einsum('ts,atb->asb', M.conj(), gA2)
        elif device_kernels:
            i = rec[1]
            if i in busy or i + 1 in busy:
                flush()
            pending.append(rec)
            busy.update((i, i + 1))
        else:
            _, i, G, A, B, A2, B2 = rec
            gA, gB = two_site_backward(A, B, G, A2, B2, cot[i], cot[i + 1], eps)
            cot[i], cot[i + 1] = gA, gB
    flush()
    if contrib:
        vals = torch.stack([v for _, v in contrib]).cpu().numpy()
        for (k, _), v in zip(contrib, vals):
            grad[k] += float(v)
    return float(e.detach()), grad
