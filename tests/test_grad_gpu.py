"""Adjoint gradients of the TFIM energy on the statevector path.

Checked against exact parameter-shift gradients of the CPU oracle at 8 qubits, and
at the 24-qubit configuration against central finite differences and parameter
shift evaluated with the GPU forward; tolerance 1e-6 as the target states
(observed errors are far below it).
"""

import math

import numpy as np
import pytest

from oracle import sv_ref

pytestmark = pytest.mark.gpu


def oracle_energy(n, theta, layers):
    return sv_ref.tfim_energy_sv(sv_ref.run_circuit(n, sv_ref.hea_gates(n, theta, layers)))


def test_adjoint_matches_oracle_parameter_shift():
    from qsim_amd.adjoint import TFIMEnergy, hardware_efficient_ansatz
    n, layers = 8, 2
    rng = np.random.default_rng(0)
    theta = rng.uniform(-math.pi, math.pi, 4 * n)
    mod = TFIMEnergy(hardware_efficient_ansatz(n, layers))
    e, g = mod.energy_and_grad(theta)
    assert abs(e - oracle_energy(n, theta, layers)) < 1e-12
    ps = np.zeros_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        ps[k] = 0.5 * (oracle_energy(n, tp, layers) - oracle_energy(n, tm, layers))
    assert np.max(np.abs(g - ps)) < 1e-10


def test_adjoint_24q_vs_fd_and_shift():
    from qsim_amd.adjoint import TFIMEnergy, hardware_efficient_ansatz
    n, layers = 24, 4
    rng = np.random.default_rng(1)
    theta = rng.uniform(-math.pi, math.pi, 2 * layers * n)
    mod = TFIMEnergy(hardware_efficient_ansatz(n, layers))
    e, g = mod.energy_and_grad(theta)
    assert g.size == 192
    sel = rng.choice(g.size, 24, replace=False)
    h = 1e-5
    for k in sel:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (mod.energy(tp) - mod.energy(tm)) / (2 * h)
        tp[k], tm[k] = theta[k] + math.pi / 2, theta[k] - math.pi / 2
        sh = 0.5 * (mod.energy(tp) - mod.energy(tm))
        assert abs(g[k] - fd) < 1e-6, (k, g[k], fd)
        assert abs(g[k] - sh) < 1e-9, (k, g[k], sh)


def test_torch_autograd_function():
    import torch
    from qsim_amd.adjoint import TFIMEnergy, hardware_efficient_ansatz, torch_energy
    n = 10
    mod = TFIMEnergy(hardware_efficient_ansatz(n, 2))
    f = torch_energy(mod)
    theta = torch.tensor(np.random.default_rng(3).uniform(-1, 1, 4 * n), dtype=torch.float64,
                         requires_grad=True)
    e = f(theta)
    e.backward()
    _, g = mod.energy_and_grad(theta.detach().numpy())
    assert torch.allclose(theta.grad, torch.as_tensor(g), atol=1e-12)
