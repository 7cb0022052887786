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

``kernel="elementwise"`` is a second, stronger tape: the same circuit and loss
with every 1-qubit gate written as slice arithmetic (y0 = m00·x0 + m01·x1,
y1 = m10·x0 + m11·x1 on the two halves of the qubit's axis), CZ as a sign
mask and the Pauli loss by flips and masks.  It avoids tensordot's tiny-GEMM
shapes (and their reduction-shaped backward) while keeping tape autograd.
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


def _apply_elementwise(mat, qubits, state):
    if len(qubits) == 1:
        q = qubits[0]
        x0, x1 = state.select(q, 0), state.select(q, 1)
        y0 = mat[0, 0] * x0 + mat[0, 1] * x1
        y1 = mat[1, 0] * x0 + mat[1, 1] * x1
        return torch.stack([y0, y1], dim=q)
    d = torch.diagonal(mat)
    if len(qubits) == 2 and torch.count_nonzero(mat - torch.diag(d)) == 0:
        shape = [1] * state.dim()
        shape[qubits[0]] = 2
        shape[qubits[1]] = 2
        if qubits[0] < qubits[1]:
            return state * d.reshape(2, 2).reshape(shape)
        return state * d.reshape(2, 2).t().reshape(shape)
    return op_action(mat, qubits, state)


def _energy_elementwise(state, terms):
    n = state.dim()
    e = torch.zeros((), dtype=torch.float32, device=state.device)
    prob = (state.conj() * state).real
    for c, ps in terms:
        right_flip = [q for q, ch in enumerate(ps) if ch in 'xy']
        zs = [q for q, ch in enumerate(ps) if ch in 'zy']
        if any(ch == 'y' for ch in ps):
            raise NotImplementedError("Y terms use the tensordot tape")
        if right_flip:
            r = torch.flip(state, right_flip)
            val = state.conj() * r
            for q in zs:
                sh = [1] * n
                sh[q] = 2
                val = val * torch.tensor([1.0, -1.0], device=state.device).reshape(sh)
            e = e + float(c) * val.real.sum()
        else:
            val = prob
            for q in zs:
                sh = [1] * n
                sh[q] = 2
                val = val * torch.tensor([1.0, -1.0], device=state.device).reshape(sh)
            e = e + float(c) * val.sum()
    return e


def tape_run(gates, n, device, dtype=torch.complex64, init=None):
    """The reference torch backend's run, restated: ψ from |0…0⟩ (or ``init``)
    through every gate in order, each matrix cast to ``dtype`` on ``device``
    (statevector_sim.py:19-24) and contracted with ``op_action`` (:124-140).
    Flat result, differentiable by tape autograd through the matrices."""
    if init is None:
        state = torch.zeros([2] * n, dtype=dtype, device=device)
        state.view(-1)[0] = 1.0
    else:
        state = torch.as_tensor(init).reshape([2] * n).to(dtype=dtype, device=device)
    for g in gates:
        m = g.matrix if torch.is_tensor(g.matrix) else torch.as_tensor(g.matrix)
        state = op_action(m.to(dtype=dtype, device=device), list(g.indices), state)
    return state.reshape(-1)


def tape_energy(circ, theta, terms, device, kernel="tensordot"):
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
            if kernel == "elementwise":
                state = _apply_elementwise(mat, list(g.indices), state)
            else:
                state = op_action(mat, list(g.indices), state)
    if kernel == "elementwise":
        return _energy_elementwise(state, terms)
    e = torch.zeros((), dtype=torch.float32, device=device)
    for c, ps in terms:
        right = state
        for q, ch in enumerate(ps):
            if ch != 'i':
                right = op_action(_pauli(ch).to(torch.complex64).to(device), [q], right)
        e = e + float(c) * torch.sum(state.conj() * right).real
    return e


def tape_energy_and_grad(circ, theta, terms, device, kernel="tensordot"):
    """Value and gradient by ``loss.backward()``.  ``theta`` on the CPU is the
    reference's setup (host Parameter, host matrix_exp, one small host->device
    copy per gate); ``theta`` on the GPU keeps the whole tape on the device,
    the strongest form of the tape baseline."""
    th = theta.detach().clone().requires_grad_(True)
    e = tape_energy(circ, th, terms, device, kernel)
    e.backward()
    return e.detach(), th.grad.detach()
