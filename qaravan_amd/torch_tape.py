"""The reference's torch backend, restated: the tape-autograd baseline.

This is what qaravan does when a parametrised circuit is differentiated
(src/qaravan/applications/compilation.py:73-133 with
``StatevectorSim(backend="torch")``): every rotation matrix is built by
``torch.linalg.matrix_exp(-1j·θ·P)`` in complex128
(core/param_gates.py:51-76), cast to complex64 on the device
(tensorQ/statevector_sim.py:19-24), and applied with a tensordot over the
gate's legs followed by a permute back into place (``op_action``,
statevector_sim.py:124-140, whose ncon call reduces to exactly that); the loss
Σ_t c_t⟨ψ|P_t|ψ⟩ applies each Pauli factor the same way
(statevector_sim.py:217-234); ``loss.backward()`` differentiates the tape.

It is used only as the comparison point for the adjoint kernels (speed and
gradient agreement); it is not a fallback of the HIP path.
"""
import torch

_PAULI = {
    'x': [[0, 1], [1, 0]],
    'y': [[0, -1j], [1j, 0]],
    'z': [[1, 0], [0, -1]],
}


def _pauli(ch, device=None):
    return torch.tensor(_PAULI[ch], dtype=torch.complex128, device=device)


def op_action(op, qubits, state):
    """Rank-n state (2,)*n; op is 2^k x 2^k with qubits[0] the row MSB."""
    k = len(qubits)
    t = op.reshape([2] * (2 * k))
    out = torch.tensordot(t, state, dims=(list(range(k, 2 * k)), list(qubits)))
    return torch.movedim(out, list(range(k)), list(qubits))


def rotation_matrix(word, theta):
    """torch.linalg.matrix_exp(-1j·θ·P) in complex128 (param_gates.py:55-76)."""
    p = _pauli(word[0], theta.device)
    for ch in word[1:]:
        p = torch.kron(p, _pauli(ch, theta.device))
    return torch.linalg.matrix_exp(-1j * theta.to(torch.complex128) * p)


def tape_energy(circ, theta, terms, device):
    """E(θ) through the restated torch backend; differentiable by autograd.
    Gates with ``param_index`` take θ[param_index] (a CPU float64 leaf, as in
    the reference where θ is a torch Parameter on the host)."""
    n = circ.num_sites
    state = torch.zeros([2] * n, dtype=torch.complex64, device=device)
    state.view(-1)[0] = 1.0
    layers = circ.layers if circ.layers is not None else circ.build().layers
    for layer in layers:
        for g in layer:
            if getattr(g, "param_index", None) is not None:
                mat = rotation_matrix(g.generator_word, theta[g.param_index])
            else:
                mat = torch.as_tensor(g.matrix, dtype=torch.complex128, device=theta.device)
            mat = mat.to(dtype=torch.complex64, device=device)
            state = op_action(mat, list(g.indices), state)
    e = torch.zeros((), dtype=torch.float32, device=device)
    for c, ps in terms:
        right = state
        for q, ch in enumerate(ps):
            if ch != 'i':
                right = op_action(_pauli(ch).to(torch.complex64).to(device), [q], right)
        e = e + float(c) * torch.sum(state.conj() * right).real
    return e


def tape_energy_and_grad(circ, theta, terms, device):
    """Value and gradient by ``loss.backward()``.  ``theta`` on the CPU is the
    reference's setup (host Parameter, host matrix_exp, one small host->device
    copy per gate); ``theta`` on the GPU keeps the whole tape on the device,
    the strongest form of the tape baseline."""
    th = theta.detach().clone().requires_grad_(True)
    e = tape_energy(circ, th, terms, device)
    e.backward()
    return e.detach(), th.grad.detach()
