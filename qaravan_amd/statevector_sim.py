"""API-compatible ``StatevectorSim`` whose gate application runs on HIP kernels.

Surface follows src/qaravan/tensorQ/statevector_sim.py:13-109 and
src/qaravan/core/base_sim.py:3-63: ``StatevectorSim(circ, init_state=None,
backend="numpy", device="cpu")``, ``run(progress_bar, debug)`` returning the
state as a rank-n array of shape (2,)*n, ``get_statevector()`` returning the
flat 2^n vector, ``state``, ``ran``, ``apply_gate``, ``normalize_state``, and
the observable helpers.  Every backend computes on the GPU; the backend string
only chooses what the caller gets back, as the reference's does
(statevector_sim.py:19-29):

* ``"numpy"`` (the default): complex128 on the GPU, results returned as host
  numpy arrays (one device-to-host copy when the state is read).
* ``"torch"``: complex64 (the reference casts to it) on the GPU, results
  returned as torch tensors on ``device`` that autograd can differentiate
  with respect to every gate matrix (and the initial state) that requires a
  gradient: the backward pass is the adjoint sweep of qaravan_amd.state_autograd,
  not a tape.
* ``"hip"``: the state stays a device tensor that the library updates in
  place (``device`` must then be a GPU; complex64 unless ``dtype`` says
  otherwise).

``dtype=`` overrides the precision of any backend.  Other differences, all
deliberate:

* ``run()`` compiles the whole layered circuit once into fused passes
  (qaravan_amd.program) instead of one contraction per gate.
* Expectation values are FP64 reductions on the device; they do not re-run the
  circuit when it has already run (the reference re-runs at :65 and :77).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .program import compile_ops, lower_circuit, lower_gate, zero_start_variant
from .utils import pretty_print_sv


def _pauli_masks(n, pauli_strings):
    """(xmask, zmask) per Pauli string; qubit q is address bit n-1-q."""
    xs, zs = [], []
    for ps in pauli_strings:
        if len(ps) != n:
            raise AssertionError("Pauli string length must match number of qubits in statevector")
        x = z = 0
        for q, ch in enumerate(ps.lower()):
            b = 1 << (n - 1 - q)
            if ch == 'x':
                x |= b
            elif ch == 'y':
                x |= b
                z |= b
            elif ch == 'z':
                z |= b
            elif ch != 'i':
                raise ValueError(f"bad Pauli letter {ch!r}")
        xs.append(x)
        zs.append(z)
    return np.array(xs, dtype=np.uint64), np.array(zs, dtype=np.uint64)


class DeviceScratch:
    """Caller-side scratch buffer for the library (it allocates nothing itself)."""

    def __init__(self, device):
        self.device = device
        self.buf = torch.empty(0, dtype=torch.uint8, device=device)

    def get(self, nbytes):
        nbytes = max(int(nbytes), 1 << 16)
        if self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self.buf


def run_program(state, prog, scratch, theta=None):
    """Run an encoded program on a flat device state tensor, in place."""
    lib = _lib.load()
    n = prog.n
    code = _lib.dtype_code(state.dtype)
    need = ctypes.c_size_t()
    _lib.check(lib.svh_program_scratch_bytes(prog.blob.ctypes.data, code, ctypes.byref(need)),
               "svh_program_scratch_bytes")
    buf = scratch.get(need.value)
    dev = prog.device_copy(state.device)
    th = None if theta is None else _lib.ptr(theta)
    sp = prog.specialized(code) if getattr(prog, "specialize", False) else None
    if sp is not None and sp.handle:
        _lib.check(lib.svh_run_program_sp(sp.handle, _lib.ptr(state), n, code, prog.blob.ctypes.data,
                                          _lib.ptr(dev), th, _lib.ptr(buf), buf.numel(),
                                          _lib.stream_ptr(state.device)), "svh_run_program_sp")
        return
    _lib.check(lib.svh_run_program(_lib.ptr(state), n, code, prog.blob.ctypes.data, _lib.ptr(dev), th,
                                   _lib.ptr(buf), buf.numel(), _lib.stream_ptr(state.device)),
               "svh_run_program")


def pauli_sum_expectation(state, n, terms, scratch):
    """Σ_t c_t ⟨ψ|P_t|ψ⟩ as a 0-d float64 device tensor (FP64 accumulation)."""
    lib = _lib.load()
    from .adjoint import real_coefficients
    coeffs = real_coefficients(terms)
    xm, zm = _pauli_masks(n, [p for _, p in terms])
    out = torch.empty(1, dtype=torch.float64, device=state.device)
    buf = scratch.get(8 * 1024 + 32 * len(terms) + 256)
    _lib.check(lib.svh_pauli_expect(_lib.ptr(state), n, _lib.dtype_code(state.dtype), len(terms),
                                    xm.ctypes.data, zm.ctypes.data, coeffs.ctypes.data, _lib.ptr(out),
                                    _lib.ptr(buf), buf.numel(), _lib.stream_ptr(state.device)),
               "svh_pauli_expect")
    return out[0]


def pauli_sum_apply(state, out, n, terms, scratch):
    """out = Σ_t c_t P_t ψ on the device."""
    lib = _lib.load()
    from .adjoint import real_coefficients
    coeffs = real_coefficients(terms)
    xm, zm = _pauli_masks(n, [p for _, p in terms])
    buf = scratch.get(32 * len(terms) + 256)
    _lib.check(lib.svh_pauli_apply(_lib.ptr(state), _lib.ptr(out), n, _lib.dtype_code(state.dtype),
                                   len(terms), xm.ctypes.data, zm.ctypes.data, coeffs.ctypes.data,
                                   _lib.ptr(buf), buf.numel(), _lib.stream_ptr(state.device)),
               "svh_pauli_apply")


def apply_matrix_inplace(flat, n, mat, indices):
    """ψ ← (M on ``indices``) ψ in place on a flat device state; ``indices``
    in gate order (first = row MSB), k ≤ 6 (statevector_sim.py:124-140)."""
    lib = _lib.load()
    mat = np.asarray(mat.detach().cpu().numpy() if torch.is_tensor(mat) else mat, dtype=np.complex128)
    k = len(indices)
    if mat.shape != (2 ** k, 2 ** k):
        raise ValueError(f"gate matrix shape {mat.shape} does not match {k} qubit(s)")
    bits = (ctypes.c_int * k)(*[n - 1 - int(q) for q in indices])
    m_dev = torch.from_numpy(np.ascontiguousarray(mat)).to(device=flat.device, dtype=flat.dtype)
    _lib.check(lib.svh_apply_dense(_lib.ptr(flat), n, _lib.dtype_code(flat.dtype), bits, k, _lib.ptr(m_dev),
                                   _lib.stream_ptr(flat.device)), "svh_apply_dense")


def op_action(op, indices, sv, local_dim=2):
    """New state (M on ``indices``) ψ, same shape and kind as ``sv``
    (statevector_sim.py:124-140), computed by one dense-gate kernel on the
    GPU: a numpy array in gives a numpy array out, a torch tensor gives a
    tensor on its device, differentiable when ``op`` or ``sv`` requires grad."""
    if local_dim != 2:
        raise NotImplementedError("the HIP statevector path supports qubits (local_dim = 2) only")
    _lib.require_gpu()
    if isinstance(indices, (int, np.integer)):
        indices = [int(indices)]
    indices = [int(q) for q in indices]
    as_numpy = not torch.is_tensor(sv)
    t = torch.as_tensor(np.asarray(sv, dtype=np.complex128)) if as_numpy else sv
    dev = t.device if t.device.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
    dtype = t.dtype if t.dtype in (torch.complex64, torch.complex128) else torch.complex128
    N = t.numel()
    n = N.bit_length() - 1
    if (1 << n) != N:
        raise ValueError(f"state has {N} amplitudes, not a power of two")
    from .gates import Gate
    from .state_autograd import differentiable_run
    g = Gate("op", indices, op)
    out = differentiable_run(n, [g], dtype, dev, None, DeviceScratch(dev), psi0=t)
    if as_numpy:
        return out.detach().cpu().numpy().reshape(np.shape(sv))
    return out.to(t.device).reshape(t.shape)


def initial_state(n, init, dtype, device):
    """Flat device state for ``init``: None = |0…0⟩, a bitstring (first
    character = qubit 0, the most significant bit), or an array/tensor of
    2^n amplitudes (statevector_sim.py:31-48)."""
    lib = _lib.load()
    code = _lib.dtype_code(dtype)
    if init is None or isinstance(init, str):
        idx = 0
        if isinstance(init, str):
            if len(init) != n:
                raise ValueError("bitstring length must equal the number of qubits")
            idx = int(init, 2)
        flat = torch.empty(2 ** n, dtype=dtype, device=device)
        _lib.check(lib.svh_init_basis(_lib.ptr(flat), n, code, idx, _lib.stream_ptr(device)), "svh_init_basis")
        return flat
    if isinstance(init, (np.ndarray, torch.Tensor)):
        sv = torch.as_tensor(init)
        if sv.numel() != 2 ** n:
            raise ValueError(f"init_state has {sv.numel()} amplitudes, expected {2 ** n}")
        return sv.detach().resolve_conj().reshape(-1).to(device=device, dtype=dtype).clone()
    raise ValueError("init_state must be a NumPy array, Torch tensor, or a bitstring")


def inner(a, b, n, scratch):
    """⟨a|b⟩ as a complex128 0-d device tensor (FP64 accumulation)."""
    lib = _lib.load()
    out = torch.empty(2, dtype=torch.float64, device=a.device)
    buf = scratch.get(16 * 1024)
    _lib.check(lib.svh_inner(_lib.ptr(a), _lib.ptr(b), n, _lib.dtype_code(a.dtype), _lib.ptr(out),
                             _lib.ptr(buf), buf.numel(), _lib.stream_ptr(a.device)), "svh_inner")
    return out.view(torch.complex128)[0]     # (re, im) read as one complex128: no kernel


_BACKENDS = ("numpy", "torch", "hip")
_DEFAULT_DTYPE = {"numpy": torch.complex128, "torch": torch.complex64, "hip": torch.complex64}


class StatevectorSim:
    """Dense statevector simulator on one MI355X (see module docstring)."""

    def __init__(self, circ, init_state=None, backend="numpy", device="cpu", dtype=None, tile=None):
        if backend not in _BACKENDS:
            raise ValueError(f"backend must be one of {_BACKENDS} (got {backend!r})")
        _lib.require_gpu()
        self.circ = circ.copy()
        self.num_sites = circ.num_sites
        self.local_dim = circ.local_dim
        if self.local_dim != 2:
            raise NotImplementedError("the HIP statevector path supports qubits (local_dim = 2) only")
        self.nm = None
        self.init_state = init_state
        self.backend = backend
        self.device = torch.device(device)
        if backend == "hip" and self.device.type != "cuda":
            raise ValueError("backend 'hip' keeps the state on a GPU device (e.g. 'cuda' or 'cuda:0')")
        # where the kernels run: the given GPU, or the current one when the
        # results are to be returned on the host
        self.compute_device = (self.device if self.device.type == "cuda"
                               else torch.device("cuda", torch.cuda.current_device()))
        self.dtype = dtype or _DEFAULT_DTYPE[backend]
        _lib.dtype_code(self.dtype)
        if tile is None:
            from .autotune import sim_tile_config
            tile = sim_tile_config(self.num_sites, self.dtype)
        self.tile = tile
        self._flat = None
        self._host = None
        self.ran = False
        self._prog = None
        self._prog0 = None     # the program with a first pass that generates |0…0⟩ (False: none)
        self._scratch = DeviceScratch(self.compute_device)

    # ------------------------------------------------------------ state
    def _set_flat(self, t):
        self._flat = t
        self._host = None

    @property
    def state(self):
        """The state as a rank-n array of shape (2,)*n: a numpy array
        ("numpy"), a torch tensor on ``device`` ("torch"), or the device
        tensor the library updates ("hip")."""
        if self._flat is None:
            return None
        shape = [self.local_dim] * self.num_sites
        if self.backend == "numpy":
            if self._host is None:
                self._host = self._flat.detach().cpu().numpy()
            return self._host.reshape(shape)
        if self.backend == "torch":
            return self._flat.to(self.device).reshape(shape)
        return self._flat.view(shape)

    @state.setter
    def state(self, value):
        if value is None:
            self._set_flat(None)
            return
        t = torch.as_tensor(value).reshape(-1)
        if not (self.backend == "torch" and t.requires_grad):
            t = t.detach().resolve_conj()
        t = t.to(device=self.compute_device, dtype=self.dtype)
        self._set_flat(t if t.requires_grad else t.contiguous().clone())

    def initialize_state(self):
        """|0…0⟩, a bitstring, or a given vector (statevector_sim.py:31-48)."""
        init = self.init_state
        if self.backend == "torch" and torch.is_tensor(init) and init.requires_grad:
            self._set_flat(init.reshape(-1).to(device=self.compute_device, dtype=self.dtype))
            return
        self._set_flat(initial_state(self.num_sites, init, self.dtype, self.compute_device))

    # ------------------------------------------------------------ gates
    def _check_no_grad(self, gates):
        if self.backend == "hip" and torch.is_grad_enabled():
            for g in gates:
                if torch.is_tensor(g.matrix) and g.matrix.requires_grad:
                    raise NotImplementedError(
                        f"gate {g.name} has a matrix that requires grad: the 'hip' backend updates the state in "
                        "place and does not differentiate; use backend='torch'")

    def program(self):
        if self._prog is None:
            circ = self.circ.build() if not self.circ.built else self.circ
            ops = lower_circuit(circ, self.num_sites)
            self._prog = compile_ops(ops, self.num_sites, self.tile)
        return self._prog

    def _program_from_zero(self):
        """program() with its first pass generating |0…0⟩ in registers (no
        initialisation sweep), or None when that pass has no compiled kernel."""
        if self._prog0 is None:
            self._prog0 = zero_start_variant(self.program(), _lib.dtype_code(self.dtype)) or False
        return self._prog0 or None

    def _run_gates(self, gates):
        """Apply ``gates`` (application order) to the current state."""
        from .state_autograd import differentiable_run
        if self.backend == "torch":
            psi0 = self._flat
            out = differentiable_run(self.num_sites, gates, self.dtype, self.compute_device, self.tile,
                                     self._scratch, psi0=psi0)
            self._set_flat(out)
            return
        self._check_no_grad(gates)
        with torch.no_grad():
            ops = []
            for g in gates:
                ops.extend(lower_gate(g, self.num_sites))
            if ops:
                run_program(self._flat, compile_ops(ops, self.num_sites, self.tile), self._scratch)
        self._host = None

    def apply_gate(self, gate):
        """Apply one gate (statevector_sim.py:50-52)."""
        if self._flat is None:
            self.initialize_state()
        self._run_gates([gate])

    def apply_reset(self, reset_op):
        self.apply_gate(reset_op)   # the lowered Reset carries its normalisation

    def run(self, progress_bar=True, debug=0):
        """Simulate the whole circuit; returns the state as shape (2,)*n
        (base_sim.py:30-53).  debug=1/2 print the state after every layer/gate."""
        circ = self.circ.build() if not self.circ.built else self.circ
        prog0 = None
        if not debug and self.backend != "torch" and self.init_state is None:
            self._check_no_grad(circ.gate_list)
            prog0 = self._program_from_zero()
        if prog0 is not None:
            # the first pass generates |0…0⟩: the buffer needs no initialisation
            self._set_flat(torch.empty(2 ** self.num_sites, dtype=self.dtype, device=self.compute_device))
        else:
            self.initialize_state()
        if debug:
            for layer in circ.layers:
                for op in layer:
                    self.apply_gate(op)
                    if debug == 2:
                        print(self._pretty())
                if debug == 1:
                    print(self._pretty())
        elif self.backend == "torch":
            gates = [g for layer in circ.layers for g in layer]
            self._run_gates(gates)
        else:
            self._check_no_grad(circ.gate_list)
            with torch.no_grad():
                prog = prog0 if prog0 is not None else self.program()
                if progress_bar:
                    from tqdm import tqdm
                    with tqdm(total=1, desc=f"{prog.n_passes} fused passes") as bar:
                        run_program(self._flat, prog, self._scratch)
                        bar.update(1)
                else:
                    run_program(self._flat, prog, self._scratch)
            self._host = None
        self.ran = True
        if circ.meas_sites is not None:
            self.measure(circ.meas_sites)      # drawn and discarded, as base_sim.py:49-50
        return self.state

    def get_statevector(self):
        """The flat 2^n state (statevector_sim.py:102-105), of the backend's kind."""
        if not self.ran:
            self.run(progress_bar=False)
        if self.backend == "numpy":
            return self.state.reshape(-1)
        if self.backend == "torch":
            return self._flat.to(self.device).reshape(-1)
        return self._flat

    def _device_state(self):
        """The current state as a flat device tensor, outside any graph."""
        if not self.ran and self._flat is None:
            self.run(progress_bar=False)
        return self._flat.detach()

    def normalize_state(self):
        if self.backend == "torch" and self._flat.requires_grad:
            self._set_flat(self._flat / torch.linalg.vector_norm(self._flat))
            return
        lib = _lib.load()
        code = _lib.dtype_code(self.dtype)
        dev = self.compute_device
        out = torch.empty(1, dtype=torch.float64, device=dev)
        buf = self._scratch.get(8 * 1024)
        st = _lib.stream_ptr(dev)
        _lib.check(lib.svh_norm2(_lib.ptr(self._flat), self.num_sites, code, _lib.ptr(out), _lib.ptr(buf),
                                 buf.numel(), st), "svh_norm2")
        _lib.check(lib.svh_scale_by_inv_norm(_lib.ptr(self._flat), self.num_sites, code, _lib.ptr(out), st),
                   "svh_scale_by_inv_norm")
        self._host = None

    # ------------------------------------------------------------ measurement
    def measure(self, meas_sites):
        """A measurement bitstring over the sorted sites (statevector_sim.py:54-57)."""
        from .measure import measure_sv
        return measure_sv(self._device_state(), meas_sites)

    def measure_and_collapse(self, meas_sites):
        """(bitstring, post-measurement state of the other sites) (statevector_sim.py:59-61);
        the state comes back in the backend's kind."""
        from .measure import measure_and_collapse_sv
        out, post = measure_and_collapse_sv(self._device_state(), meas_sites)
        if self.backend == "numpy":
            return out, post.cpu().numpy()
        if self.backend == "torch":
            return out, post.to(self.device)
        return out, post

    # ------------------------------------------------------------ observables
    def pauli_expectation(self, pauli_string):
        """Re⟨ψ|P|ψ⟩ for a string over 'ixyz' (statevector_sim.py:75-79,217-234)."""
        if not self.ran:
            self.run(progress_bar=False)
        return float(pauli_sum_expectation(self._device_state(), self.num_sites, [(1.0, pauli_string)],
                                           self._scratch).item())

    def expectation(self, terms):
        """Σ c·⟨P⟩ over (coeff, pauli string) terms, e.g. Hamiltonian.pauli_terms()."""
        if not self.ran:
            self.run(progress_bar=False)
        return float(pauli_sum_expectation(self._device_state(), self.num_sites, terms, self._scratch).item())

    def _local_ops_expectation(self, gates):
        psi = self._device_state()
        right = psi.clone()
        ops = []
        for g in gates:
            ops.extend(lower_gate(g, self.num_sites))
        run_program(right, compile_ops(ops, self.num_sites, self.tile, fuse=False), self._scratch)
        return float(inner(psi, right, self.num_sites, self._scratch).real.item())

    def one_local_expectation(self, op, site):
        """Re⟨ψ|O_site|ψ⟩ for a 2x2 O (statevector_sim.py:81-94)."""
        if not self.ran:
            self.run(progress_bar=False)
        from .gates import Gate
        return self._local_ops_expectation([Gate("O", [site], np.asarray(op))])

    def local_expectation(self, local_ops):
        """Re⟨ψ|⊗_i O_i|ψ⟩ for one 2x2 per site (statevector_sim.py:63-73)."""
        if not self.ran:
            self.run(progress_bar=False)
        from .gates import Gate
        return self._local_ops_expectation([Gate("O", [i], np.asarray(local_ops[i]))
                                            for i in range(self.num_sites)])

    def _pretty(self):
        return pretty_print_sv(self._flat.detach().cpu().numpy(), self.local_dim)

    def __str__(self):
        if not self.ran:
            self.run(progress_bar=False)
        return self._pretty()
