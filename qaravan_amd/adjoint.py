"""Adjoint differentiation of ⟨ψ(θ)|H|ψ(θ)⟩ behind ``torch.autograd.Function``.

The reference differentiates by tape autograd through its torch backend
(src/qaravan/applications/compilation.py:73-133: matrix_exp gates, tensordot per
gate, ``cost.backward()``), which stores every intermediate state: O(G·2^n)
memory.  Here the backward is the adjoint method in O(2^n):

  forward   ψ_G = U_G…U_1|0…0⟩ (fused tile passes, θ read on the device),
            λ_G = Hψ_G (Pauli-sum kernel), E = Re⟨ψ_G|λ_G⟩ (FP64 reduction)
  backward  for k = G…1:  dE/dθ_k += 2·Im⟨λ_k|P_k|ψ_k⟩   (gate k = exp(−iθ_k P_k))
                          ψ_{k-1} = U_k† ψ_k,  λ_{k-1} = U_k† λ_k
            done by one program of reversed, daggered ops with SVH_OPF_GRAD on
            the rotations; the ⟨λ|Pψ⟩ sums are accumulated in FP64 per
            workgroup and reduced in a fixed order.

Rotations are qaravan's RX/RY/RZ/RXX/RYY/RZZ (U = exp(−iθP), no ½,
core/param_gates.py:51-103); a gate takes its angle from θ[g.param_index].
"""
from dataclasses import replace

import numpy as np
import torch

from . import _lib
from .program import OPF_GRAD, TileConfig, compile_ops, lower_circuit, zero_start_variant
from .statevector_sim import DeviceScratch, run_program


def _adjoint_cfg(n, dtype, cfg):
    """The adjoint pass holds ψ and λ: cap the tile so both fit in LDS and the
    workgroup so the op-interpreting kernel fits in registers (c64: ≤ 512
    threads, c128: ≤ 256; one more thread bit with 3 register bits).  Five
    register bits are forward-only."""
    rb = min(cfg.rb, 4)
    tb_max = (9 if dtype == torch.complex64 else 8) + (4 - rb)
    if cfg.generic:
        # at most 512 threads, and ψ and λ tiles within 160 KiB of LDS
        tb_max = min(9, (13 if dtype == torch.complex64 else 12) - rb)
    tb = min(cfg.tb, tb_max, n - rb)
    return replace(cfg, tb=tb, rb=rb, prefetch=False)


class ParamProgram:
    """Forward and adjoint programs of a parametrised circuit plus its loss."""

    def __init__(self, circ, terms, n_params, cfg=None, adj_cfg=None, dtype=torch.complex64, ops=None, n=None,
                 start_zero=False):
        """``circ``: a qaravan-style Circuit whose rotations carry ``param_index``;
        or ``circ=None`` with ``ops`` (elementary ops, e.g. workloads.hea_ops) and ``n``."""
        self.n = circ.num_sites if circ is not None else int(n)
        if self.n < 8:
            raise NotImplementedError("the adjoint path needs n >= 8 qubits")
        self.terms = list(terms)
        self.n_params = int(n_params)
        self.dtype = dtype
        cfg = cfg or TileConfig()
        adj_cfg = adj_cfg or cfg
        if ops is None:
            ops = lower_circuit(circ if circ.built else circ.build(), self.n, param_mode=True)
        for op in ops:
            if op.kind in ('dense', 'norm'):
                raise NotImplementedError("gates on > 2 qubits and resets are not differentiable here")
            if op.kind in ('rot1', 'rot2') and not 0 <= op.param < self.n_params:
                raise ValueError(f"param index {op.param} outside [0, {self.n_params})")
        # a remapped forward program leaves ψ in its own layout (no permutation
        # back): λ = Hψ is formed in that layout with relabelled Pauli masks,
        # ⟨ψ|λ⟩ does not depend on it, and the adjoint program starts from it
        self.fwd = compile_ops(ops, self.n, replace(cfg, tb=min(cfg.tb, self.n - cfg.rb), prefetch=False),
                               n_params=self.n_params, keep_layout=True)
        self.layout = self.fwd.end_pos    # physical bit of logical bit q after the forward program (None: natural)
        # start_zero: the forward's first pass generates |0…0⟩ itself when it
        # has a compiled kernel (fwd0; the caller skips the initialisation)
        self.fwd0 = zero_start_variant(self.fwd, _lib.dtype_code(dtype)) if start_zero else None
        back = []
        for op in reversed(ops):
            d = op.dagger()
            if d.kind in ('rot1', 'rot2'):
                d.flags |= OPF_GRAD
            back.append(d)
        self.adj = compile_ops(back, self.n, _adjoint_cfg(self.n, dtype, adj_cfg), grad=True,
                               n_params=self.n_params, pos0=self.layout)
        self.n_gates = sum(1 for _ in ops)


def real_coefficients(terms):
    """The real coefficients of a Pauli sum; a coefficient with an imaginary
    part (beyond rounding, 1e-12·|c|) makes the sum non-Hermitian, whose
    expectation ⟨ψ|H|ψ⟩ the loss here does not define, so it is refused."""
    out = []
    for c, p in terms:
        c = complex(c)
        if abs(c.imag) > 1e-12 * max(abs(c), 1e-300):
            raise ValueError(f"term {p!r} has complex coefficient {c}: the Pauli sum must be Hermitian")
        out.append(c.real)
    return np.array(out, dtype=np.float64)


def _relabel_mask(m, layout):
    out = 0
    for b in range(len(layout)):
        if (int(m) >> b) & 1:
            out |= 1 << layout[b]
    return out


def term_table(n, terms, layout=None):
    """Device-layout Pauli-sum table: (xmask, zmask, Re, Im of c·i^{ny}) per term.
    ``layout``: the state's address-bit layout (physical bit of logical bit b),
    whose bits the masks are moved to (default: the natural order)."""
    dt = np.dtype([('x', '<u8'), ('z', '<u8'), ('cre', '<f8'), ('cim', '<f8')])
    from .statevector_sim import _pauli_masks
    xm, zm = _pauli_masks(n, [p for _, p in terms])
    if layout is not None:
        xm = np.array([_relabel_mask(x, layout) for x in xm], dtype=np.uint64)
        zm = np.array([_relabel_mask(z, layout) for z in zm], dtype=np.uint64)
    cs = real_coefficients(terms)
    t = np.zeros(len(terms), dtype=dt)
    ph = [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)]
    for i, c in enumerate(cs):
        ny = bin(int(xm[i]) & int(zm[i])).count("1") & 3
        t[i] = (xm[i], zm[i], c * ph[ny][0], c * ph[ny][1])
    return t


def _theta64(theta, device):
    return theta.detach().to(device=device, dtype=torch.float64).contiguous()


def forward_state(pp, theta, device, scratch):
    """ψ(θ), λ = Hψ (device tensors, both in the forward program's final layout
    ``pp.layout``; None: the natural order), the float64 θ and the energy
    ⟨ψ|H|ψ⟩ (0-d float64 device tensor, FP64 sums in the same sweep as λ)."""
    lib = _lib.load()
    n = pp.n
    th = _theta64(theta, device)
    psi = torch.empty(2 ** n, dtype=pp.dtype, device=device)
    if pp.fwd0 is not None:
        run_program(psi, pp.fwd0, scratch, th)       # its first pass generates |0…0⟩
    else:
        _lib.check(lib.svh_init_basis(_lib.ptr(psi), n, _lib.dtype_code(pp.dtype), 0, _lib.stream_ptr(device)),
                   "svh_init_basis")
        run_program(psi, pp.fwd, scratch, th)
    lam = torch.empty_like(psi)
    key = ("terms", str(device))
    if key not in pp.__dict__:
        pp.__dict__[key] = torch.from_numpy(term_table(n, pp.terms, pp.layout).view(np.uint8).copy()).to(device)
    tt = pp.__dict__[key]
    out = torch.empty(2, dtype=torch.float64, device=device)
    buf = scratch.get(16 * 1024)
    _lib.check(lib.svh_pauli_apply_inner(_lib.ptr(psi), _lib.ptr(lam), n, _lib.dtype_code(pp.dtype), len(pp.terms),
                                         _lib.ptr(tt), _lib.ptr(out), _lib.ptr(buf), buf.numel(),
                                         _lib.stream_ptr(device)), "svh_pauli_apply_inner")
    return psi, lam, th, out[0]


def adjoint_sweep(pp, psi, lam, th, scratch):
    """dE/dθ (float64 device tensor); consumes ψ and λ (walked back in place)."""
    import ctypes
    lib = _lib.load()
    n = pp.n
    grad = torch.zeros(pp.n_params, dtype=torch.float64, device=psi.device)
    code = _lib.dtype_code(psi.dtype)
    need = ctypes.c_size_t()
    _lib.check(lib.svh_program_scratch_bytes(pp.adj.blob.ctypes.data, code, ctypes.byref(need)),
               "svh_program_scratch_bytes")
    buf = scratch.get(need.value)
    dev = pp.adj.device_copy(psi.device)
    sp = pp.adj.specialized(code, dual=True) if pp.adj.specialize else None
    if sp is not None and sp.handle:
        _lib.check(lib.svh_run_adjoint_sp(sp.handle, _lib.ptr(psi), _lib.ptr(lam), n, code, pp.adj.blob.ctypes.data,
                                          _lib.ptr(dev), _lib.ptr(th), _lib.ptr(grad), _lib.ptr(buf), buf.numel(),
                                          _lib.stream_ptr(psi.device)), "svh_run_adjoint_sp")
        return grad
    _lib.check(lib.svh_run_adjoint(_lib.ptr(psi), _lib.ptr(lam), n, code, pp.adj.blob.ctypes.data,
                                   _lib.ptr(dev), _lib.ptr(th), _lib.ptr(grad), _lib.ptr(buf), buf.numel(),
                                   _lib.stream_ptr(psi.device)), "svh_run_adjoint")
    return grad


class AdjointEnergy(torch.autograd.Function):
    """E(θ) = ⟨ψ(θ)|H|ψ(θ)⟩ with an adjoint backward; saves only ψ and λ."""

    @staticmethod
    def forward(ctx, theta, pp, scratch):
        psi, lam, th, e = forward_state(pp, theta, theta.device, scratch)
        ctx.pp, ctx.psi, ctx.lam, ctx.th, ctx.scratch = pp, psi, lam, th, scratch
        ctx.theta_dtype = theta.dtype
        return e.to(torch.float64)

    @staticmethod
    def backward(ctx, g_e):
        if ctx.psi is None:
            # the sweep walks ψ and λ back in place: they are gone after one backward
            raise RuntimeError("AdjointEnergy: backward ran twice through the same graph (retain_graph=True); "
                               "the adjoint sweep consumes the saved state, so evaluate the energy again instead")
        grad = adjoint_sweep(ctx.pp, ctx.psi, ctx.lam, ctx.th, ctx.scratch)
        ctx.psi = ctx.lam = None
        return (g_e * grad).to(ctx.theta_dtype), None, None


class AdjointModel:
    """Callable loss E(θ) for one parametrised circuit, differentiable by torch."""

    def __init__(self, circ, terms, n_params, device="cuda", dtype=torch.complex64, cfg=None):
        _lib.require_gpu()
        self.device = torch.device(device)
        self.pp = ParamProgram(circ, terms, n_params, cfg=cfg, dtype=dtype, start_zero=True)
        self.scratch = DeviceScratch(self.device)

    def __call__(self, theta):
        return AdjointEnergy.apply(theta, self.pp, self.scratch)

    def energy_and_grad(self, theta):
        """(E, dE/dθ) without building an autograd graph."""
        psi, lam, th, e = forward_state(self.pp, theta, self.device, self.scratch)
        g = adjoint_sweep(self.pp, psi, lam, th, self.scratch)
        return e, g

    def capture(self, theta):
        """Record one energy+gradient evaluation in a HIP graph.  The sequence
        reads θ from a static device buffer and copies nothing from the host,
        so ``replay(θ')`` re-runs it for new angles as one launch."""
        th = theta.detach().to(device=self.device, dtype=torch.float64).contiguous().clone()
        for _ in range(2):  # warm-up: compile/load kernels, size the scratch buffer
            self.energy_and_grad(th)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                e, grad = self.energy_and_grad(th)
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        self._graph = (g, th, e, grad)
        return g

    def replay(self, theta):
        """(E, dE/dθ) for new θ through the captured graph (capture() first)."""
        g, th, e, grad = self._graph
        th.copy_(theta.detach().to(device=self.device, dtype=torch.float64))
        g.replay()
        return e, grad
