"""Differentiable statevector runs: gate matrices (and the initial state) -> ψ.

The reference differentiates a circuit by running its torch backend under tape
autograd: every gate matrix is a torch tensor in the graph (e.g. built by
``kak_unitary`` from a ``torch.nn.Parameter``, src/qaravan/core/param_gates.py:
247-260, or ``torch.linalg.matrix_exp`` of an angle, :51-76), it is cast to
complex64 on the device (tensorQ/statevector_sim.py:19-24) and contracted into
the state one tensordot at a time (:124-140); ``cost.backward()`` then walks
the tape (applications/compilation.py:73-90,123; examples/
state_prep_with_ad.ipynb).  The tape keeps every intermediate state.

Here the forward pass is the fused HIP program (one compiled run of the whole
circuit) and the backward pass is the adjoint sweep of the same gate list,
O(2^n) memory whatever the depth.  For ψ_G = M_G…M_1 ψ_0 and an upstream
gradient g = ∂L/∂ψ_G (torch's convention for complex tensors: the vector-
Jacobian product of y = M·x is M^H·g for x and g·x^H for M):

    λ ← g
    for k = G … 1:   ψ ← M_k^† ψ                       (ψ_{k-1}, uncomputed: M_k unitary)
                     ∂L/∂M_k[i, j] = Σ_r λ(i, r)·conj(ψ(j, r))   (partial overlap over the
                                                       gate's qubits, FP64 sums)
                     λ ← M_k^† λ
    ∂L/∂ψ_0 = λ

Uncomputing ψ needs unitary gates; a non-unitary gate (a Reset, a projector)
in a differentiated run raises instead of returning a wrong gradient.  The
sweep stops uncomputing ψ below the earliest gate that needs a gradient, and
stops altogether there when ψ_0 needs none.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .program import compile_ops, lower_gate, lower_matrix_gate


class GradPlan:
    """The gate list of one differentiable run, in application order.

    ``gates``: list of (gate, input index or None); a gate with an input index
    takes its matrix from the autograd inputs, the others from ``gate.matrix``.
    ``init``: how ψ_0 is formed when no state tensor is passed (None, a
    bitstring or an array), as ``StatevectorSim.initialize_state``."""

    def __init__(self, n, gates, dtype, device, tile, scratch, init=None):
        self.n = int(n)
        self.gates = list(gates)
        self.dtype = dtype
        self.device = device
        self.tile = tile
        self.scratch = scratch
        self.init = init

    # ------------------------------------------------------------ helpers
    def _matrix(self, gate, idx, mats):
        m = mats[idx] if idx is not None else gate.matrix
        if torch.is_tensor(m):
            m = m.detach().resolve_conj().cpu().numpy()
        return np.asarray(m, dtype=np.complex128)

    def _initial(self, psi0):
        from .statevector_sim import initial_state
        if psi0 is not None:
            t = psi0.detach().resolve_conj().reshape(-1).to(device=self.device, dtype=self.dtype).clone()
            if t.numel() != 2 ** self.n:
                raise ValueError(f"initial state has {t.numel()} amplitudes, expected {2 ** self.n}")
            return t
        return initial_state(self.n, self.init, self.dtype, self.device)

    # ------------------------------------------------------------ forward
    def forward(self, psi0, mats):
        from .statevector_sim import run_program
        psi = self._initial(psi0)
        ops = []
        for g, idx in self.gates:
            if idx is None:
                ops.extend(lower_gate(g, self.n))
            else:
                ops.extend(lower_matrix_gate(self._matrix(g, idx, mats), list(g.indices), self.n, g.name))
        if ops:
            run_program(psi, compile_ops(ops, self.n, self.tile), self.scratch)
        return psi

    # ------------------------------------------------------------ backward
    def backward(self, out, grad, mats, need_psi0):
        """(∂L/∂ψ_0 or None, [∂L/∂M for each input matrix])."""
        lib = _lib.load()
        n = self.n
        code = _lib.dtype_code(self.dtype)
        st = _lib.stream_ptr(self.device)
        want = [idx is not None and mats[idx].requires_grad for _, idx in self.gates]
        first = next((k for k, w in enumerate(want) if w), len(self.gates))
        stop = 0 if need_psi0 else first
        # every gate the sweep passes: M^† in the state precision, one upload
        mats_np, offs, total = [], [], 0
        for k in range(stop, len(self.gates)):
            g, idx = self.gates[k]
            m = self._matrix(g, idx, mats)
            d = m.shape[0]
            if m.shape != (d, d) or d != 2 ** len(g.indices):
                raise ValueError(f"gate {g.name} matrix shape {m.shape} does not match its {len(g.indices)} qubit(s)")
            if len(g.indices) > 6:
                raise NotImplementedError("gates on more than 6 qubits are not differentiable here")
            if g.name == "Reset" or not np.allclose(m.conj().T @ m, np.eye(d), atol=1e-6):
                raise NotImplementedError(
                    f"gate {g.name} on {list(g.indices)} is not unitary: the adjoint sweep uncomputes "
                    "the state with M^dagger, so non-unitary gates (Reset, projectors) are not differentiable")
            mats_np.append(m.conj().T)
            offs.append(total)
            total += d * d
        host = np.zeros(max(total, 1), dtype=np.complex128)
        for m, o in zip(mats_np, offs):
            host[o:o + m.size] = m.reshape(-1)
        ddev = torch.from_numpy(host).to(device=self.device, dtype=self.dtype)
        es = ddev.element_size()
        # ∂L/∂M blocks, complex128, one buffer
        goffs, gtot = {}, 0
        for k in range(first, len(self.gates)):
            if want[k]:
                goffs[k] = gtot
                gtot += 4 ** len(self.gates[k][0].indices)
        gbuf = torch.zeros(max(gtot, 1), dtype=torch.complex128, device=self.device)
        lam = grad.detach().resolve_conj().reshape(-1).to(device=self.device, dtype=self.dtype).contiguous().clone()
        psi = out.detach().resolve_conj().reshape(-1).to(device=self.device, dtype=self.dtype).contiguous().clone()
        max_k = max((len(self.gates[k][0].indices) for k in goffs), default=0)
        obuf = self.scratch.get(lib.svh_measure_scratch_bytes(_lib.SVH_MEAS_OVERLAP, n, max_k)) if goffs else None
        for k in reversed(range(stop, len(self.gates))):
            g, _ = self.gates[k]
            kq = len(g.indices)
            bits = (ctypes.c_int * kq)(*[n - 1 - int(q) for q in g.indices])
            mptr = ctypes.c_void_p(ddev.data_ptr() + es * offs[k - stop])
            if k >= first:       # ψ_{k-1}: still needed for this or an earlier gradient
                _lib.check(lib.svh_apply_dense(_lib.ptr(psi), n, code, bits, kq, mptr, st), "svh_apply_dense")
            if want[k]:
                optr = ctypes.c_void_p(gbuf.data_ptr() + 16 * goffs[k])
                _lib.check(lib.svh_partial_overlap(_lib.ptr(lam), _lib.ptr(psi), n, code, bits, kq, optr,
                                                   _lib.ptr(obuf), obuf.numel(), st), "svh_partial_overlap")
            _lib.check(lib.svh_apply_dense(_lib.ptr(lam), n, code, bits, kq, mptr, st), "svh_apply_dense")
        gm = [None] * len(mats)
        host_g = None
        for k, o in goffs.items():
            g, idx = self.gates[k]
            d = 2 ** len(g.indices)
            m = mats[idx]
            blk = gbuf[o:o + d * d]
            if m.device.type == "cpu":
                if host_g is None:
                    host_g = gbuf.cpu()
                blk = host_g[o:o + d * d]
            blk = blk.reshape(d, d).to(device=m.device)
            blk = blk if m.is_complex() else blk.real
            blk = blk.to(m.dtype)
            gm[idx] = blk if gm[idx] is None else gm[idx] + blk
        return (lam if need_psi0 else None), gm


class RunFunction(torch.autograd.Function):
    """ψ = run(ψ_0, M_1, M_2, …) with the adjoint backward of ``GradPlan``."""

    @staticmethod
    def forward(ctx, plan, psi0, *mats):
        out = plan.forward(psi0, mats)
        ctx.plan = plan
        ctx.mats = mats
        ctx.psi0_meta = None if psi0 is None else (psi0.shape, psi0.dtype, psi0.device)
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, grad):
        (out,) = ctx.saved_tensors
        need_psi0 = bool(ctx.needs_input_grad[1])
        gpsi, gm = ctx.plan.backward(out, grad, ctx.mats, need_psi0)
        if gpsi is not None:
            shape, dt, dev = ctx.psi0_meta
            gpsi = gpsi.reshape(shape).to(device=dev, dtype=dt if dt.is_complex else gpsi.dtype)
            if not dt.is_complex:
                gpsi = gpsi.real.to(dt)
        return (None, gpsi, *gm)


def differentiable_run(n, gates, dtype, device, tile, scratch, init=None, psi0=None):
    """Run ``gates`` (application order) from ψ_0 and return the flat device
    state; it carries an autograd graph when a gate matrix or ``psi0`` is a
    tensor that requires grad (and grad mode is on)."""
    mats, spec = [], []
    for g in gates:
        m = g.matrix
        if torch.is_tensor(m) and m.requires_grad:
            spec.append((g, len(mats)))
            mats.append(m)
        else:
            spec.append((g, None))
    plan = GradPlan(n, spec, dtype, device, tile, scratch, init)
    if torch.is_grad_enabled() and (mats or (psi0 is not None and psi0.requires_grad)):
        return RunFunction.apply(plan, psi0, *mats)
    with torch.no_grad():
        return plan.forward(psi0, mats)
