"""The tape-autograd baselines (restated torch backend) agree with the oracle:
energies and gradients on CPU tensors, both gate kernels."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd.torch_tape import tape_energy_and_grad
from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms


@pytest.mark.parametrize("kernel", ["tensordot", "elementwise"])
def test_tape_matches_oracle(kernel):
    n, layers = 6, 2
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    e, g = tape_energy_and_grad(circ, torch.tensor(theta), tfim_terms(n), torch.device("cpu"), kernel)
    spec = []
    for l in range(layers):
        for k, w in enumerate("yz"):
            for q in range(n):
                spec.append((w, [q], (2 * l + k) * n + q, None))
        for i in range(n - 1):
            spec.append((None, [i, i + 1], None, np.diag([1, 1, 1, -1]).astype(complex)))
    er, gr = O.adjoint_gradient(spec, theta, n, O.tfi_terms(n, -1.0, -1.0))
    assert abs(e.item() - er) < 1e-5
    assert np.linalg.norm(g.numpy() - gr) / np.linalg.norm(gr) < 1e-5
