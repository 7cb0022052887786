"""λ = Hψ for Pauli sums of every length modulo the kernels' term grouping
(4 terms per step in the single-circuit kernel, 2 in the batched one), with
X, Y, Z and mixed strings, checked through the adjoint energy and gradient
against the oracle (oracle/sv_oracle.py adjoint_gradient)."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms

pytestmark = pytest.mark.gpu


def extra_terms(n, k, seed):
    """k random Pauli strings over i, x, y, z with coefficients in (-1, 1)."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(-1, 1)), "".join(rng.choice(list("ixyz"), size=n))) for _ in range(k)]


def spec_of(circ):
    spec = []
    for g in circ.gate_list:
        if getattr(g, "param_index", None) is not None:
            spec.append((g.generator_word, g.indices, g.param_index, None))
        else:
            spec.append((None, g.indices, None, np.asarray(g.matrix, dtype=complex)))
    return spec


@pytest.mark.parametrize("extra", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [8, 10, 11])
def test_adjoint_pauli_sum_lengths(gpu, n, extra):
    from qaravan_amd.adjoint import AdjointModel
    theta = hea_theta(n, 2)
    circ = hea_circuit(n, 2, theta)
    terms = tfim_terms(n) + extra_terms(n, extra, 100 + 4 * n + extra)
    print(f"n={n} terms={len(terms)} (mod 4 = {len(terms) % 4})", flush=True)
    m = AdjointModel(circ, terms, len(theta), device=gpu, dtype=torch.complex128)
    e, g = m.energy_and_grad(torch.tensor(theta, device=gpu))
    er, gr = O.adjoint_gradient(spec_of(circ), theta, n, terms)
    assert abs(e.item() - er) < 1e-11
    assert np.linalg.norm(g.cpu().numpy() - gr) / np.linalg.norm(gr) < 1e-10


@pytest.mark.parametrize("extra", [0, 1])
def test_batch_pauli_sum_lengths(gpu, extra):
    from qaravan_amd.batch import BatchedAdjoint
    circs, terms, thetas = [], [], []
    for c, n in enumerate((10, 11, 10)):
        th = hea_theta(n, 1 + c)
        circs.append(hea_circuit(n, 1 + c, th))
        terms.append(tfim_terms(n) + extra_terms(n, extra + c, 200 + 4 * n + extra + c))
        thetas.append(th)
    print("term counts", [len(t) for t in terms], flush=True)
    b = BatchedAdjoint(circs, terms, [len(t) for t in thetas], device=gpu, dtype=torch.complex128,
                       specialize=True, lds_split=1)
    b.set_theta(thetas)
    e, _ = b.run()
    grads = b.gradients()
    for c, circ in enumerate(circs):
        er, gr = O.adjoint_gradient(spec_of(circ), thetas[c], circ.num_sites, terms[c])
        assert abs(e[c].item() - er) < 1e-11
        assert np.linalg.norm(grads[c] - gr) / np.linalg.norm(gr) < 1e-10
