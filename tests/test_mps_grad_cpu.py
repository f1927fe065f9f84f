"""CPU checks of the MPS backward (qsim_amd.mps_grad): truncated-SVD backward and the
whole reverse sweep against central finite differences and the statevector oracle.

The forward here is a small torch restatement of the two-site update (mps_core.py:277-350:
SVD, keep = clip(#{S > cutoff S0}, 1, chi_max), S_t = S/||S_k||, A' = U sqrt(S_t),
B' = sqrt(S_t) V^H) producing the same tape the GPU driver records.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import sv_ref
from qsim_amd import mps_grad
from qsim_amd.adjoint import hardware_efficient_ansatz
from qsim_amd.gates import matrix_2q


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_svd_backward_matches_finite_differences():
    rng = np.random.default_rng(11)
    m, n, k = 6, 5, 3
    th0 = _rand(rng, m, k) @ _rand(rng, k, n)              # exact rank k
    W1, W2, W3 = _rand(rng, m, n), rng.standard_normal((m, k)), rng.standard_normal((k, n))

    def parts(th):
        U, S, Vh = np.linalg.svd(th, full_matrices=False)
        return U[:, :k], S[:k], Vh[:k].conj().T

    def loss_np(th):
        U, S, V = parts(th)
        A2 = U * np.sqrt(S)
        B2 = np.sqrt(S)[:, None] * V.conj().T
        return (np.real(np.sum(np.conj(W1) * (A2 @ B2))) + np.sum(W2 * np.abs(A2) ** 2) +
                np.sum(W3 * np.abs(B2) ** 2))

    U, S, V = (torch.as_tensor(x) for x in parts(th0))
    A2 = (U * torch.sqrt(S).to(U.dtype)).detach().requires_grad_(True)
    B2 = (torch.sqrt(S).to(U.dtype)[:, None] * V.conj().T).detach().requires_grad_(True)
    L = (torch.real(torch.sum(torch.as_tensor(W1).conj() * (A2 @ B2))) +
         torch.sum(torch.as_tensor(W2) * A2.abs() ** 2) + torch.sum(torch.as_tensor(W3) * B2.abs() ** 2))
    gA2, gB2 = torch.autograd.grad(L, (A2, B2))
    D = torch.sqrt(S).to(U.dtype)
    gU = gA2 * D[None, :]
    gV = gB2.conj().T * D[None, :]
    gS = torch.sqrt(S) / (2 * S) * (torch.real(torch.diagonal(U.conj().T @ gA2)) +
                                     torch.real(torch.diagonal(gB2 @ V)))
    g = mps_grad.svd_backward(U, S, V, gU, gS, gV).numpy()
    h = 1e-6
    for (i, j) in [(0, 0), (2, 3), (5, 4), (3, 1)]:
        for unit in (1.0, 1j):
            E = np.zeros((m, n), dtype=complex)
            E[i, j] = unit * h
            fd = (loss_np(th0 + E) - loss_np(th0 - E)) / (2 * h)
            an = np.real(np.conj(g[i, j]) * unit)
            assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (i, j, unit, fd, an)


def _torch_forward(nq, theta, layers, chi_max=64, cutoff=1e-12):
    """Torch restatement of the GPU forward with the driver's tape."""
    c = hardware_efficient_ansatz(nq, layers)
    sites = [torch.zeros(1, 2, 1, dtype=torch.complex128) for _ in range(nq)]
    for s in sites:
        s[0, 0, 0] = 1.0
    tape = []
    for blk in c.blocks:
        if blk.kind == 'rot':
            for name, q, k in blk.ops:
                tape.append(('1q', q, name, k, sites[q].clone()))
                M = torch.as_tensor(mps_grad.rotation(name, theta[k]))
                sites[q] = torch.einsum('st,atb->asb', M, sites[q])
        else:
            for g in blk.ops:
                i = min(g[1])
                G = torch.as_tensor(matrix_2q(g[0]))
                A, B = sites[i], sites[i + 1]
                cl, cr = A.shape[0], B.shape[2]
                T = torch.einsum('apb,bqd->apqd', A, B).reshape(cl, 4, cr)
                th = torch.einsum('st,atd->asd', G, T).reshape(2 * cl, 2 * cr)
                U, S, Vh = torch.linalg.svd(th, full_matrices=False)
                keep = int(max(1, min(chi_max, int((S > cutoff * S[0]).sum()))))
                St = S[:keep] / torch.linalg.norm(S[:keep])
                A2 = (U[:, :keep] * torch.sqrt(St).to(U.dtype)).reshape(cl, 2, keep)
                B2 = (torch.sqrt(St).to(U.dtype)[:, None] * Vh[:keep]).reshape(keep, 2, cr)
                tape.append(('2q', i, G, A.clone(), B.clone(), A2.clone(), B2.clone()))
                sites[i], sites[i + 1] = A2, B2
    return sites, tape, c.nparams


def _sv_energy(nq, theta, layers):
    st = sv_ref.run_circuit(nq, sv_ref.hea_gates(nq, theta, layers))
    return sv_ref.tfim_energy_sv(st)


@pytest.mark.parametrize('nq,layers,seed', [(4, 2, 0), (6, 3, 1)])
def test_reverse_sweep_matches_fd_and_statevector(nq, layers, seed):
    rng = np.random.default_rng(seed)
    npar = 2 * nq * layers
    theta = rng.uniform(-math.pi, math.pi, npar)
    sites, tape, _ = _torch_forward(nq, theta, layers)
    e, g = mps_grad.energy_and_grad_from_tape(sites, tape, theta, npar)
    assert abs(e - _sv_energy(nq, theta, layers)) < 1e-10
    # parameter shift on the statevector oracle (exact for exp(-i theta P / 2))
    for k in rng.choice(npar, size=min(npar, 8), replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        ps = 0.5 * (_sv_energy(nq, tp, layers) - _sv_energy(nq, tm, layers))
        assert abs(g[k] - ps) < 1e-8, (k, g[k], ps)
    # central differences of the MPS forward itself
    for k in rng.choice(npar, size=3, replace=False):
        h = 1e-5
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        ep = mps_grad.tfim_energy(_torch_forward(nq, tp, layers)[0]).item()
        em = mps_grad.tfim_energy(_torch_forward(nq, tm, layers)[0]).item()
        assert abs(g[k] - (ep - em) / (2 * h)) < 1e-6, k


def test_broadened_kernel_is_bounded_at_degeneracy():
    # exactly degenerate singular values: the broadened F keeps the result finite
    U = torch.eye(3, dtype=torch.complex128)[:, :2]
    V = torch.eye(3, dtype=torch.complex128)[:, :2]
    S = torch.tensor([0.5, 0.5], dtype=torch.float64)
    gU = torch.ones(3, 2, dtype=torch.complex128)
    gV = torch.ones(3, 2, dtype=torch.complex128)
    gS = torch.zeros(2, dtype=torch.float64)
    g = mps_grad.svd_backward(U, S, V, gU, gS, gV)
    assert torch.isfinite(g.real).all() and torch.isfinite(g.imag).all()


def test_reverse_sweep_runs_with_grad_mode_off():
    """torch.autograd.Function.backward runs with grad mode off; the reverse sweep's
    energy gradient must still build its own graph there."""
    import torch
    rng = np.random.default_rng(5)
    dims = [1, 2, 3, 2, 1]
    sites = [torch.tensor(rng.standard_normal((dims[i], 2, dims[i + 1]))
                          + 1j * rng.standard_normal((dims[i], 2, dims[i + 1])),
                          dtype=torch.complex128) for i in range(4)]
    e_on, g_on = mps_grad.energy_and_grad_from_tape(sites, [], np.zeros(0), 0)
    with torch.no_grad():
        e_off, g_off = mps_grad.energy_and_grad_from_tape(sites, [], np.zeros(0), 0)
    assert e_on == e_off and g_on.shape == g_off.shape == (0,)
    assert math.isfinite(e_on)
