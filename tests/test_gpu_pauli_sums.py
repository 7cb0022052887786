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


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,extra", [(9, 3), (12, 6)])
def test_pauli_apply_inner_matches_apply_then_inner(gpu, dtype, n, extra):
    """svh_pauli_apply_inner: the same λ = Hψ as svh_pauli_apply_table, and
    ⟨ψ|λ⟩ (re, im) within the state precision of svh_inner's, against the
    oracle's ⟨ψ|H|ψ⟩ on a random (unnormalised) state."""
    from qaravan_amd import _lib
    from qaravan_amd.adjoint import term_table
    from qaravan_amd.statevector_sim import inner
    from qaravan_amd.statevector_sim import DeviceScratch
    rng = np.random.default_rng(n + extra)
    sv = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    terms = tfim_terms(n) + extra_terms(n, extra, 7 * n)
    print(f"pauli_apply_inner n={n} terms={len(terms)} {dtype}", flush=True)
    psi = torch.tensor(sv, dtype=dtype, device=gpu)
    tt = torch.from_numpy(term_table(n, terms).view(np.uint8).copy()).to(gpu)
    lib = _lib.load()
    code = _lib.dtype_code(dtype)
    st = _lib.stream_ptr(gpu)
    lam_a, lam_b = torch.empty_like(psi), torch.empty_like(psi)
    _lib.check(lib.svh_pauli_apply_table(_lib.ptr(psi), _lib.ptr(lam_a), n, code, len(terms), _lib.ptr(tt), st),
               "svh_pauli_apply_table")
    out = torch.empty(2, dtype=torch.float64, device=gpu)
    scratch = DeviceScratch(gpu)
    buf = scratch.get(16 * 1024)
    _lib.check(lib.svh_pauli_apply_inner(_lib.ptr(psi), _lib.ptr(lam_b), n, code, len(terms), _lib.ptr(tt),
                                         _lib.ptr(out), _lib.ptr(buf), buf.numel(), st), "svh_pauli_apply_inner")
    assert torch.equal(lam_a, lam_b)
    sep = inner(psi, lam_a, n, DeviceScratch(gpu)).item()
    fused = complex(out[0].item(), out[1].item())
    ref = np.vdot(sv.astype(np.complex128) if dtype == torch.complex128 else psi.cpu().numpy().astype(np.complex128),
                  O.apply_pauli_sum(psi.cpu().numpy().astype(np.complex128), terms))
    scale = np.linalg.norm(sv) ** 2 * sum(abs(c) for c, _ in terms)
    tol = (1e-13 if dtype == torch.complex128 else 1e-6) * scale
    assert abs(fused - ref) < tol, (fused, ref)
    assert abs(sep - ref) < tol, (sep, ref)
