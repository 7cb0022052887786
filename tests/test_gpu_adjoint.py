"""Adjoint gradients on the GPU vs the oracle (parameter shift is exact for
exp(-iθP)), vs central differences, and vs the torch tape path.

Tolerances: complex128 ‖g − g_ref‖/‖g_ref‖ ≤ 1e-10; complex64 ≤ 1e-5."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import program as P
from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms

pytestmark = pytest.mark.gpu


def hea_spec(n, layers):
    spec = []
    for l in range(layers):
        for k, w in enumerate("yz"):
            for q in range(n):
                spec.append((w, [q], (2 * l + k) * n + q, None))
        for i in range(n - 1):
            spec.append((None, [i, i + 1], None, np.diag([1, 1, 1, -1]).astype(complex)))
    return spec


@pytest.mark.parametrize("n,layers,tb,spz,rb,split", [(8, 2, 4, False, 4, 0), (10, 3, 6, False, 4, 0),
                                                    (12, 2, 8, False, 4, 0), (10, 3, 6, True, 4, 0),
                                                    (12, 2, 5, True, 4, 0), (11, 3, 7, True, 3, 0),
                                                    (12, 4, 9, True, 3, 0), (10, 3, 6, True, 4, 1),
                                                    (12, 4, 9, True, 3, 1), (12, 2, 7, True, 4, 2),
                                                    (14, 4, 8, True, 3, "remap"), (13, 3, 7, True, 4, "remap")])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_adjoint_vs_oracle(gpu, n, layers, tb, spz, rb, split, dtype):
    from qaravan_amd.adjoint import AdjointModel
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    terms = tfim_terms(n)
    print(f"adjoint n={n} layers={layers} tb={tb} rb={rb} specialize={spz} lds_split={split} {dtype}", flush=True)
    model = AdjointModel(circ, terms, len(theta), device=gpu, dtype=dtype,
                         cfg=P.TileConfig(tb=tb, low_bits=6 if split != "remap" else 4, max_cost=48, specialize=spz,
                                          rb=rb, lds_split=0 if split == "remap" else split,
                                          remap=split == "remap", io_lane_bits=4 if split == "remap" else 6))
    if split == "remap":
        # the forward ends in its own layout; λ and the adjoint sweep start from it
        assert model.pp.layout is not None and model.pp.fwd.final_pos is None
    th = torch.tensor(theta, device=gpu, requires_grad=True)
    e = model(th)
    e.backward()
    e_ref, g_ref = O.adjoint_gradient(hea_spec(n, layers), theta, n, O.tfi_terms(n, -1.0, -1.0))
    g = th.grad.cpu().numpy()
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    tol = 1e-10 if dtype == torch.complex128 else 1e-5
    print(f"  |E-E_ref|={abs(e.item() - e_ref):.2e} rel grad err={rel:.2e}", flush=True)
    assert abs(e.item() - e_ref) < (1e-11 if dtype == torch.complex128 else 1e-5)
    assert rel < tol
    gps = O.param_shift_gradient(hea_spec(n, layers), theta, n, O.tfi_terms(n, -1.0, -1.0), idx=range(0, len(theta), 7))
    # each component is within the relative gradient tolerance of ‖g_ref‖ (parameter shift is exact)
    assert np.allclose(g[::7], gps, rtol=0, atol=tol * np.linalg.norm(g_ref))


def test_adjoint_energy_and_grad_matches_autograd(gpu):
    from qaravan_amd.adjoint import AdjointModel
    n, layers = 10, 2
    theta = hea_theta(n, layers)
    model = AdjointModel(hea_circuit(n, layers, theta), tfim_terms(n), len(theta), device=gpu,
                         dtype=torch.complex128)
    th = torch.tensor(theta, device=gpu, requires_grad=True)
    e = model(th)
    e.backward()
    e2, g2 = model.energy_and_grad(th.detach())
    assert abs(e.item() - e2.item()) < 1e-13
    assert torch.allclose(th.grad, g2, atol=1e-13)
    # a second call reuses the compiled programs and gives the same numbers
    e3, g3 = model.energy_and_grad(th.detach())
    assert torch.equal(g2, g3)
    # the sweep consumes the saved state: a second backward through one graph raises
    th2 = torch.tensor(theta, device=gpu, requires_grad=True)
    e4 = model(th2)
    e4.backward(retain_graph=True)
    with pytest.raises(RuntimeError, match="backward ran twice"):
        e4.backward()


def test_adjoint_vs_torch_tape(gpu):
    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.torch_tape import tape_energy_and_grad
    n, layers = 10, 3
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    terms = tfim_terms(n)
    model = AdjointModel(circ, terms, len(theta), device=gpu, dtype=torch.complex64)
    e, g = model.energy_and_grad(torch.tensor(theta, device=gpu))
    e_t, g_t = tape_energy_and_grad(circ, torch.tensor(theta, dtype=torch.float64), terms, gpu)
    assert abs(e.item() - e_t.item()) < 1e-5
    assert np.linalg.norm(g.cpu().numpy() - g_t.numpy()) / np.linalg.norm(g_t.numpy()) < 1e-5


def test_adjoint_20q_c64_vs_c128_and_fd(gpu):
    """The 20-qubit, 16-layer ansatz (640 parameters): c64 gradient within 1e-5
    of the c128 gradient; c128 gradient vs central differences (ε = 1e-3,
    complex128 forward runs on the GPU) on a subset, bias sin(2ε)/(2ε) removed."""
    from qaravan_amd.adjoint import AdjointModel, forward_state
    from qaravan_amd.statevector_sim import inner
    n, layers = 20, 16
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    terms = tfim_terms(n)
    m64 = AdjointModel(circ, terms, len(theta), device=gpu, dtype=torch.complex64)
    m128 = AdjointModel(circ, terms, len(theta), device=gpu, dtype=torch.complex128)
    th = torch.tensor(theta, device=gpu)
    # the tuned configuration (specialised kernels, 3 or 4 register bits) agrees too
    from qaravan_amd.autotune import best_adjoint_config
    mt = AdjointModel(circ, terms, len(theta), device=gpu, dtype=torch.complex64,
                      cfg=best_adjoint_config(n, torch.complex64))
    _, gt = mt.energy_and_grad(th)
    e64, g64 = m64.energy_and_grad(th)
    e128, g128 = m128.energy_and_grad(th)
    g64, g128 = g64.cpu().numpy(), g128.cpu().numpy()
    rel = np.linalg.norm(g64 - g128) / np.linalg.norm(g128)
    rel_t = np.linalg.norm(gt.cpu().numpy() - g128) / np.linalg.norm(g128)
    print(f"20q: E64={e64.item():.9f} E128={e128.item():.12f} rel grad err c64={rel:.2e} tuned={rel_t:.2e}",
          flush=True)
    assert rel < 1e-5 and rel_t < 1e-5
    eps = 1e-3
    bias = np.sin(2 * eps) / (2 * eps)

    def energy(t):
        psi, lam, _, e = forward_state(m128.pp, t, gpu, m128.scratch)
        # the fused energy equals ⟨ψ|λ⟩ taken separately
        assert abs(e.item() - inner(psi, lam, n, m128.scratch).real.item()) < 1e-12
        return e.item()
    for i in (0, 33, 321, 639):
        tp, tm = th.clone(), th.clone()
        tp[i] += eps
        tm[i] -= eps
        fd = (energy(tp) - energy(tm)) / (2 * eps) / bias
        assert abs(fd - g128[i]) < 1e-8 * max(1.0, abs(g128[i])), (i, fd, g128[i])


@pytest.mark.parametrize("spz", [False, True])
def test_adjoint_graph_replay_matches_eager(gpu, spz):
    """Captured energy+gradient replayed for several θ, with eager work in
    between, equals the eager evaluation each time."""
    from qaravan_amd.adjoint import AdjointModel
    n, layers = 12, 3
    theta = hea_theta(n, layers)
    m = AdjointModel(hea_circuit(n, layers, theta), tfim_terms(n), len(theta), device=gpu,
                     dtype=torch.complex64, cfg=P.TileConfig(tb=6, low_bits=4, max_cost=1e9, specialize=spz))
    th = torch.tensor(theta, device=gpu)
    m.capture(th)
    for k in range(3):
        t = th + 0.05 * k
        e_ref, g_ref = m.energy_and_grad(t)
        e_ref, g_ref = e_ref.clone(), g_ref.clone()
        _ = torch.ones(100, device=gpu).sum().item()
        e, g = m.replay(t)
        torch.cuda.synchronize()
        assert torch.allclose(e, e_ref, atol=1e-9)
        assert torch.allclose(g, g_ref, atol=1e-9)
