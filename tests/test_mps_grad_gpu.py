"""MPS gradients on the GPU engine (qsim_amd.mps_grad): the reverse sweep through the
recorded two-site updates against the statevector oracle (parameter shift) at small n, and
against central finite differences of the GPU MPS forward on the 200-qubit chi_max 32
hardware-efficient ansatz."""

import math

import numpy as np
import pytest

from oracle import sv_ref

pytestmark = pytest.mark.gpu


def _sv_energy(nq, theta, layers):
    return sv_ref.tfim_energy_sv(sv_ref.run_circuit(nq, sv_ref.hea_gates(nq, theta, layers)))


def test_mps_gradient_matches_statevector_parameter_shift():
    from qsim_amd.adjoint import hardware_efficient_ansatz
    from qsim_amd.mps_grad import MPSEnergyGrad
    nq, layers = 10, 4
    rng = np.random.default_rng(5)
    c = hardware_efficient_ansatz(nq, layers)
    theta = rng.uniform(-math.pi, math.pi, c.nparams)
    eng = MPSEnergyGrad(c, chi_max=32, cutoff=1e-12)
    e, g = eng.energy_and_grad(theta)
    assert abs(e - _sv_energy(nq, theta, layers)) < 1e-10
    for k in rng.choice(c.nparams, size=12, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        ps = 0.5 * (_sv_energy(nq, tp, layers) - _sv_energy(nq, tm, layers))
        assert abs(g[k] - ps) < 1e-8, (k, g[k], ps)


def test_mps_gradient_200q_chi32_matches_finite_differences():
    from qsim_amd.adjoint import hardware_efficient_ansatz
    from qsim_amd.mps_grad import MPSEnergyGrad
    nq, layers = 200, 4
    rng = np.random.default_rng(9)
    c = hardware_efficient_ansatz(nq, layers)
    theta = rng.uniform(-math.pi, math.pi, c.nparams)
    eng = MPSEnergyGrad(c, chi_max=32, cutoff=1e-12)
    e, g = eng.energy_and_grad(theta)
    assert abs(e - eng.energy(theta)) < 1e-12 * max(1.0, abs(e))
    assert g.shape == (1600,) and np.all(np.isfinite(g))
    h = 1e-5
    for k in [0, 399, 803, 1234, 1599]:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (eng.energy(tp) - eng.energy(tm)) / (2 * h)
        assert abs(g[k] - fd) < 1e-6, (k, g[k], fd)
    # parameter shift on the MPS forward itself (exact for exp(-i theta P / 2))
    for k in [17, 1000]:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        ps = 0.5 * (eng.energy(tp) - eng.energy(tm))
        assert abs(g[k] - ps) < 1e-8, (k, g[k], ps)


def test_mps_gradient_through_torch_autograd():
    import torch
    from qsim_amd.adjoint import hardware_efficient_ansatz
    from qsim_amd.mps_grad import MPSEnergyGrad, torch_energy
    nq, layers = 12, 3
    rng = np.random.default_rng(17)
    c = hardware_efficient_ansatz(nq, layers)
    theta0 = rng.uniform(-math.pi, math.pi, c.nparams)
    eng = MPSEnergyGrad(c, chi_max=16, cutoff=1e-12)
    e_ref, g_ref = eng.energy_and_grad(theta0)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    e = torch_energy(eng)(theta)
    (2.0 * e).backward()
    assert abs(e.item() - e_ref) < 1e-12
    assert np.max(np.abs(theta.grad.numpy() - 2.0 * g_ref)) < 1e-12


def test_device_backward_kernel_matches_torch_reverse_sweep():
    """The hand-written batched two-site backward (qs_mps_backward_batch) against the same
    formulas as torch tensor ops, on one forward tape: the 16-qubit ansatz at chi_max 8
    (truncating) and 32 (exact), and the 200-qubit chi_max 32 case's first gradients."""
    from qsim_amd.adjoint import hardware_efficient_ansatz
    from qsim_amd.mps_grad import MPSEnergyGrad, energy_and_grad_from_tape
    rng = np.random.default_rng(23)
    for nq, chi in ((16, 8), (16, 32), (60, 32)):
        c = hardware_efficient_ansatz(nq, 4)
        theta = rng.uniform(-math.pi, math.pi, c.nparams)
        eng = MPSEnergyGrad(c, chi_max=chi, cutoff=1e-12)
        finals, tape = eng._run(theta, record=True)
        e1, g1 = energy_and_grad_from_tape(finals, tape, theta, c.nparams, device_kernels=True)
        e2, g2 = energy_and_grad_from_tape(finals, tape, theta, c.nparams, device_kernels=False)
        assert e1 == e2
        scale = max(1.0, float(np.max(np.abs(g2))))
        assert np.max(np.abs(g1 - g2)) < 1e-10 * scale, (nq, chi, np.max(np.abs(g1 - g2)))
