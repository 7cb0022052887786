"""Batched forward + adjoint (many circuits, one launch sequence / one HIP
graph) vs per-circuit adjoint runs and the oracle."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms

pytestmark = pytest.mark.gpu


# 3 register bits per thread: switched off in BatchedAdjoint (the 512-circuit
# benchmark batch faulted the GPU once; the cause is not found yet)
RB3 = pytest.param("rb3", marks=pytest.mark.skip(reason="3-register-bit batches are switched off: the "
                                                         "512-circuit batch faulted the GPU once, cause not found"))


def make_batch(nc, seed, nmin=10, nmax=14, lmax=5):
    rng = np.random.default_rng(seed)
    circs, terms, thetas = [], [], []
    for c in range(nc):
        n = int(rng.integers(nmin, nmax + 1))
        layers = int(rng.integers(1, lmax + 1))
        th = rng.uniform(-np.pi, np.pi, size=2 * layers * n)
        circs.append(hea_circuit(n, layers, th))
        terms.append(tfim_terms(n))
        thetas.append(th)
    return circs, terms, thetas


@pytest.mark.parametrize("spz", [False, True, "split", "remap", RB3])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_batch_matches_single_circuit_adjoint(gpu, dtype, spz):
    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.batch import BatchedAdjoint
    circs, terms, thetas = make_batch(9, 5)
    b = BatchedAdjoint(circs, terms, [len(t) for t in thetas], device=gpu, dtype=dtype, specialize=bool(spz),
                       lds_split=1 if spz in ("split", "remap") else 0, remap=spz == "remap",
                       rb=3 if spz == "rb3" else 4)
    b.set_theta(thetas)
    e, _ = b.run()
    e = e.cpu().numpy()
    grads = b.gradients()
    tol = 1e-10 if dtype == torch.complex128 else 2e-5
    for c in range(len(circs)):
        m = AdjointModel(circs[c], terms[c], len(thetas[c]), device=gpu, dtype=dtype)
        ec, gc = m.energy_and_grad(torch.tensor(thetas[c], device=gpu))
        gc = gc.cpu().numpy()
        print(f"circuit {c}: n={circs[c].num_sites} |dE|={abs(e[c] - ec.item()):.2e} "
              f"rel dg={np.linalg.norm(grads[c] - gc) / np.linalg.norm(gc):.2e}", flush=True)
        assert abs(e[c] - ec.item()) < tol
        assert np.linalg.norm(grads[c] - gc) / np.linalg.norm(gc) < tol


@pytest.mark.parametrize("spz", [False, True, "split", "remap", RB3])
def test_batch_vs_oracle_c128(gpu, spz):
    from qaravan_amd.batch import BatchedAdjoint
    circs, terms, thetas = make_batch(4, 11, nmin=10, nmax=11, lmax=2)
    b = BatchedAdjoint(circs, terms, [len(t) for t in thetas], device=gpu, dtype=torch.complex128,
                       specialize=bool(spz), lds_split=1 if spz in ("split", "remap") else 0, remap=spz == "remap",
                       rb=3 if spz == "rb3" else 4)
    b.set_theta(thetas)
    e, _ = b.run()
    grads = b.gradients()
    for c, circ in enumerate(circs):
        n = circ.num_sites
        spec = []
        for g in circ.gate_list:
            if getattr(g, "param_index", None) is not None:
                spec.append((g.generator_word, g.indices, g.param_index, None))
            else:
                spec.append((None, g.indices, None, np.asarray(g.matrix, dtype=complex)))
        er, gr = O.adjoint_gradient(spec, thetas[c], n, O.tfi_terms(n, -1.0, -1.0))
        assert abs(e[c].item() - er) < 1e-11
        assert np.linalg.norm(grads[c] - gr) / np.linalg.norm(gr) < 1e-10


@pytest.mark.parametrize("spz", [False, True, "split", "remap", RB3])
def test_batch_graph_replays_match_eager(gpu, spz):
    from qaravan_amd.batch import BatchedAdjoint
    circs, terms, thetas = make_batch(7, 23)
    b = BatchedAdjoint(circs, terms, [len(t) for t in thetas], device=gpu, dtype=torch.complex64,
                       specialize=bool(spz), lds_split=1 if spz in ("split", "remap") else 0, remap=spz == "remap",
                       rb=3 if spz == "rb3" else 4)
    b.set_theta(thetas)
    e0, g0 = b.run()
    e0, g0 = e0.clone(), g0.clone()
    b.capture()
    for rep in range(3):
        torch.cuda.synchronize()
        x = torch.ones(1000, device=gpu).sum()  # eager work between replays
        e, g = b.run()
        torch.cuda.synchronize()
        assert torch.allclose(e, e0, atol=1e-6), rep
        assert torch.allclose(g, g0, atol=1e-6), rep
        assert x.item() == 1000
    # new θ through the same graph
    th2 = [t + 0.1 for t in thetas]
    b.set_theta(th2)
    e2, g2 = b.run()
    e2, g2 = e2.clone(), g2.clone()
    b.graph = None
    e3, g3 = b.run()
    assert torch.allclose(e2, e3, atol=1e-6) and torch.allclose(g2, g3, atol=1e-6)
    lat = b.latencies_us()
    assert np.all(lat > 0)
