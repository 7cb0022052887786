"""Parity at the benchmark's full sizes, through the paths users call.

* The 28-qubit, 40-layer brickwork circuit through ``StatevectorSim``'s default
  (tuned, compiled, remapped) configuration in complex64, against a complex128
  run through the op-interpreting kernels (a different code path):
  ‖ψ64 − ψ128‖₂ ≤ 1e-5, both norms kept (1e-6 / 1e-12).
* The same circuit followed by its inverse returns |0…0⟩ in both precisions
  (|⟨0|ψ⟩ − 1| ≤ 1e-5 complex64, ≤ 1e-11 complex128).
* An independent check of the 28-qubit state itself: after the first 6 layers
  the reduced density matrix of a pair of qubits depends only on the gates in
  the pair's backward light cone (at most 14 qubits), which the oracle
  simulates exactly; the GPU's 28-qubit rdm must match it (≤ 1e-12 complex128,
  ≤ 1e-6 complex64).
* C2 at 30 qubits: the complex64 MFMA blocks against complex128 through the
  dense kernel (‖Δψ‖ ≤ 1e-5), and light-cone rdms of a 3-layer circuit against
  the oracle.
* The 512-circuit batch (10..22 qubits, 4..32 layers) in one persistent launch,
  complex64: energies and gradients of 8 seeded circuits (up to 16 qubits) against the oracle's
  adjoint (relative gradient error ≤ 1e-5, |ΔE| ≤ 1e-5), and of the two largest
  circuits against the complex128 single-circuit adjoint on the GPU (which the
  adjoint tests pin to the oracle at smaller sizes).
"""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128


def _light_cone(gates, sites):
    """Gates that can influence the reduced state of ``sites`` (backward light
    cone, application order) and the qubits they touch (sorted)."""
    cone, kept = set(sites), []
    for g in reversed(gates):
        if cone & set(g.indices):
            kept.append(g)
            cone |= set(g.indices)
    return kept[::-1], sorted(cone)


class _Relabelled:
    def __init__(self, g, where):
        self.name, self.matrix = g.name, g.matrix
        self.indices = [where[q] for q in g.indices]


def _run_28q(gpu, circ, dtype):
    from qaravan_amd import StatevectorSim
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=dtype)
    sim.run(progress_bar=False)
    return sim


def test_brickwork_28q_tuned_c64_vs_c128_and_inverse(gpu):
    from qaravan_amd import Circuit
    from qaravan_amd.workloads import brickwork_circuit
    n = 28
    circ = brickwork_circuit(n, 40)
    print("28q x 40 brickwork: tuned complex64 vs op-interpreting complex128", flush=True)
    s64 = _run_28q(gpu, circ, C64)
    prog = s64.program()
    assert prog.specialize and prog.final_pos is not None, "the default 28-qubit configuration is the tuned one"
    a = s64.get_statevector()
    s128 = _run_28q(gpu, circ, C128)
    b = s128.get_statevector()
    err = torch.linalg.vector_norm(a.to(C128) - b).item()
    print(f"  |psi64 - psi128| = {err:.3e}", flush=True)
    assert abs(torch.linalg.vector_norm(b).item() - 1) < 1e-12
    assert abs(torch.linalg.vector_norm(a.to(C128)).item() - 1) < 1e-6
    assert err <= 1e-5
    del s64, s128, a, b
    torch.cuda.empty_cache()
    inv = Circuit(circ.gate_list + [g.dag() for g in circ.gate_list[::-1]], n=n)
    for dtype, tol in ((C64, 1e-5), (C128, 1e-11)):
        print(f"28q U then U^dagger, {dtype}", flush=True)
        s = _run_28q(gpu, inv, dtype)
        z0 = complex(s.get_statevector()[0].item())
        print(f"  <0|psi> = {z0}", flush=True)
        assert abs(z0 - 1) <= tol
        del s
        torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype,tol", [(C128, 1e-12), (C64, 1e-6)])
def test_brickwork_28q_light_cone_rdms_vs_oracle(gpu, dtype, tol):
    from qaravan_amd.measure import rdm_from_sv
    from qaravan_amd.workloads import brickwork_circuit
    n, layers = 28, 6
    circ = brickwork_circuit(n, layers)
    print(f"28q x {layers} brickwork light-cone rdms, {dtype}", flush=True)
    sim = _run_28q(gpu, circ, dtype)
    psi = sim.get_statevector()
    for pair in ([0, 1], [13, 14], [26, 27]):
        kept, cone = _light_cone(circ.gate_list, pair)
        assert len(cone) <= 16
        where = {q: i for i, q in enumerate(cone)}
        ref = O.run([_Relabelled(g, where) for g in kept], len(cone))
        r_ref = O.rdm_from_sv(ref, [where[q] for q in pair])
        r_gpu = rdm_from_sv(psi, pair).cpu().numpy()
        err = np.abs(r_gpu - r_ref).max()
        print(f"  pair {pair}: cone of {len(cone)} qubits, {len(kept)} gates, max |drho| = {err:.3e}", flush=True)
        assert err <= tol


def test_haar_blocks_30q_c64_vs_c128_and_light_cone(gpu):
    """C2 at its full size (30 qubits, 24 layers of Haar 5-qubit blocks): the
    complex64 MFMA path against complex128 through the dense kernel (another
    code path), ‖Δψ‖ ≤ 1e-5; and after 3 layers the 30-qubit rdms of three
    qubit pairs against the oracle's simulation of their light cones."""
    from qaravan_amd.measure import rdm_from_sv
    from qaravan_amd.workloads import haar_block_circuit
    n = 30
    circ = haar_block_circuit(n, 24)
    print("30q x 24 Haar 5-qubit blocks: complex64 (MFMA) vs complex128 (dense kernel)", flush=True)
    a = _run_28q(gpu, circ, C64).get_statevector()
    torch.cuda.empty_cache()
    b = _run_28q(gpu, circ, C128).get_statevector()
    err = torch.linalg.vector_norm(a.to(C128) - b).item()
    print(f"  |psi64 - psi128| = {err:.3e}", flush=True)
    assert abs(torch.linalg.vector_norm(b).item() - 1) < 1e-12
    assert err <= 1e-5
    del a, b
    torch.cuda.empty_cache()
    circ = haar_block_circuit(n, 3)
    for dtype, tol in ((C128, 1e-12), (C64, 1e-6)):
        psi = _run_28q(gpu, circ, dtype).get_statevector()
        for pair in ([0, 1], [14, 15], [28, 29]):
            kept, cone = _light_cone(circ.gate_list, pair)
            assert len(cone) <= 16
            where = {q: i for i, q in enumerate(cone)}
            ref = O.run([_Relabelled(g, where) for g in kept], len(cone))
            err = np.abs(rdm_from_sv(psi, pair).cpu().numpy() - O.rdm_from_sv(ref, [where[q] for q in pair])).max()
            print(f"  {dtype} pair {pair}: cone of {len(cone)} qubits, max |drho| = {err:.3e}", flush=True)
            assert err <= tol
        del psi
        torch.cuda.empty_cache()


def _hea_spec(n, layers):
    spec = []
    for l in range(layers):
        for k, w in enumerate("yz"):
            spec += [(w, [q], (2 * l + k) * n + q, None) for q in range(n)]
        spec += [(None, [i, i + 1], None, np.diag([1, 1, 1, -1]).astype(complex)) for i in range(n - 1)]
    return spec


def test_persistent_512_batch_sample_vs_oracle(gpu):
    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.persistent import BatchPlan, PersistentBatch
    from qaravan_amd.workloads import batch_specs, hea_circuit, hea_ops, tfim_terms
    ns, ls, thetas, _ = batch_specs(512, 20260004)
    print("512-circuit batch, one persistent launch, complex64", flush=True)
    plan = BatchPlan(ns.tolist(), [hea_ops(int(n), int(l)) for n, l in zip(ns, ls)],
                     [tfim_terms(int(n)) for n in ns], [len(t) for t in thetas],
                     keys=list(zip(ns.tolist(), ls.tolist())))
    b = PersistentBatch(None, None, None, None, device=gpu, plan=plan)
    b.set_theta(thetas)
    energies, _ = b.run()
    torch.cuda.synchronize(gpu)
    b.check()
    energies = energies.cpu().numpy()
    grads = b.gradients()
    rng = np.random.default_rng(20261204)
    # circuits the oracle finishes in seconds: up to 16 qubits at 8 layers, 14 qubits at 32
    small = [i for i in range(512) if 2 ** int(ns[i]) * int(ls[i]) <= 2 ** 19]
    sample = sorted(int(i) for i in rng.choice(small, size=8, replace=False))
    for c in sample:
        n, L = int(ns[c]), int(ls[c])
        e_ref, g_ref = O.adjoint_gradient(_hea_spec(n, L), thetas[c], n, O.tfi_terms(n, -1.0, -1.0))
        rel = np.linalg.norm(grads[c] - g_ref) / np.linalg.norm(g_ref)
        print(f"  circuit {c}: {n}q x {L}: |dE| = {abs(energies[c] - e_ref):.2e}, rel grad err = {rel:.2e}",
              flush=True)
        assert abs(energies[c] - e_ref) <= 1e-5
        assert rel <= 1e-5
    for c in [int(i) for i in np.argsort(ns * 100 + ls)[-2:]]:
        n, L = int(ns[c]), int(ls[c])
        m = AdjointModel(hea_circuit(n, L, thetas[c]), tfim_terms(n), len(thetas[c]), device=gpu, dtype=C128)
        e128, g128 = m.energy_and_grad(torch.tensor(thetas[c], device=gpu))
        g128 = g128.cpu().numpy()
        rel = np.linalg.norm(grads[c] - g128) / np.linalg.norm(g128)
        print(f"  circuit {c}: {n}q x {L}: |dE| vs c128 = {abs(energies[c] - e128.item()):.2e}, "
              f"rel grad err = {rel:.2e}", flush=True)
        assert abs(energies[c] - e128.item()) <= 1e-5
        assert rel <= 1e-5
        del m
