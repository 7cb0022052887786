"""API-compatible ``StatevectorSim`` whose gate application runs on HIP kernels.

Surface follows src/qaravan/tensorQ/statevector_sim.py:13-109 and
src/qaravan/core/base_sim.py:3-63: ``StatevectorSim(circ, init_state=None,
backend=..., device=...)``, ``run(progress_bar, debug)`` returning the state as
a rank-n tensor of shape (2,)*n, ``get_statevector()`` returning the flat 2^n
vector, ``state``, ``ran``, ``apply_gate``, ``normalize_state``, and the
observable helpers.  Differences, all deliberate:

* ``backend="hip"`` (the default and only value): the state is a torch tensor
  on ``device`` (complex64 by default, ``dtype=torch.complex128`` optional) that
  is updated IN PLACE by the library; the reference builds a new array per gate
  (statevector_sim.py:139-140).
* ``run()`` compiles the whole layered circuit once into fused passes
  (qaravan_amd.program) instead of one contraction per gate.
* Expectation values are FP64 reductions on the device; they do not re-run the
  circuit when it has already run (the reference re-runs at :65 and :77).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .program import TileConfig, compile_ops, lower_circuit, lower_gate
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
    coeffs = np.array([float(np.real(c)) for c, _ in terms], dtype=np.float64)
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
    coeffs = np.array([float(np.real(c)) for c, _ in terms], dtype=np.float64)
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
    """New state (M on ``indices``) ψ, same shape as ``sv`` (statevector_sim.py:
    124-140): a device copy of ``sv`` updated by one dense-gate kernel."""
    if local_dim != 2:
        raise NotImplementedError("the HIP statevector path supports qubits (local_dim = 2) only")
    t = sv if torch.is_tensor(sv) else torch.as_tensor(np.asarray(sv, dtype=np.complex128))
    if t.device.type != "cuda":
        _lib.require_gpu()
        t = t.to("cuda")
    flat = t.reshape(-1).clone()
    n = flat.numel().bit_length() - 1
    if isinstance(indices, (int, np.integer)):
        indices = [int(indices)]
    apply_matrix_inplace(flat, n, op, list(indices))
    return flat.view(t.shape)


def inner(a, b, n, scratch):
    """⟨a|b⟩ as a complex128 0-d device tensor (FP64 accumulation)."""
    lib = _lib.load()
    out = torch.empty(2, dtype=torch.float64, device=a.device)
    buf = scratch.get(16 * 1024)
    _lib.check(lib.svh_inner(_lib.ptr(a), _lib.ptr(b), n, _lib.dtype_code(a.dtype), _lib.ptr(out),
                             _lib.ptr(buf), buf.numel(), _lib.stream_ptr(a.device)), "svh_inner")
    return torch.complex(out[0], out[1])


class StatevectorSim:
    """Dense statevector simulator on one MI355X (see module docstring)."""

    def __init__(self, circ, init_state=None, backend="hip", device="cuda", dtype=torch.complex64,
                 tile=None):
        if backend != "hip":
            raise ValueError(f"backend must be 'hip' (got {backend!r}); this package runs the statevector "
                             "path only on the GPU")
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
        if self.device.type != "cuda":
            raise ValueError("device must be a GPU device (e.g. 'cuda' or 'cuda:0')")
        self.dtype = dtype
        _lib.dtype_code(dtype)
        self.tile = tile or TileConfig()
        self._flat = None
        self.ran = False
        self._prog = None
        self._scratch = DeviceScratch(self.device)

    # ------------------------------------------------------------ state
    @property
    def state(self):
        if self._flat is None:
            return None
        return self._flat.view([self.local_dim] * self.num_sites)

    @state.setter
    def state(self, value):
        if value is None:
            self._flat = None
            return
        t = torch.as_tensor(value).to(device=self.device, dtype=self.dtype).reshape(-1).contiguous()
        self._flat = t

    def initialize_state(self):
        """|0…0⟩, a bitstring, or a given vector (statevector_sim.py:31-48)."""
        n = self.num_sites
        lib = _lib.load()
        code = _lib.dtype_code(self.dtype)
        if self.init_state is None or isinstance(self.init_state, str):
            idx = 0
            if isinstance(self.init_state, str):
                if len(self.init_state) != n:
                    raise ValueError("bitstring length must equal the number of qubits")
                idx = int(self.init_state, 2)
            if self._flat is None or self._flat.numel() != 2 ** n:
                self._flat = torch.empty(2 ** n, dtype=self.dtype, device=self.device)
            _lib.check(lib.svh_init_basis(_lib.ptr(self._flat), n, code, idx, _lib.stream_ptr(self.device)),
                       "svh_init_basis")
        elif isinstance(self.init_state, (np.ndarray, torch.Tensor)):
            sv = torch.as_tensor(self.init_state)
            if sv.numel() != 2 ** n:
                raise ValueError(f"init_state has {sv.numel()} amplitudes, expected {2 ** n}")
            self._flat = sv.reshape(-1).to(device=self.device, dtype=self.dtype).clone()
        else:
            raise ValueError("init_state must be a NumPy array, Torch tensor, or a bitstring")

    # ------------------------------------------------------------ gates
    def program(self):
        if self._prog is None:
            circ = self.circ.build() if not self.circ.built else self.circ
            ops = lower_circuit(circ, self.num_sites)
            self._prog = compile_ops(ops, self.num_sites, self.tile)
        return self._prog

    def apply_gate(self, gate):
        """Apply one gate in place (statevector_sim.py:50-52)."""
        if self._flat is None:
            self.initialize_state()
        prog = compile_ops(lower_gate(gate, self.num_sites), self.num_sites, self.tile)
        run_program(self._flat, prog, self._scratch)

    def apply_reset(self, reset_op):
        self.apply_gate(reset_op)   # the lowered Reset carries its normalisation

    def run(self, progress_bar=True, debug=0):
        """Simulate the whole circuit; returns the state as shape (2,)*n
        (base_sim.py:30-53).  debug=1/2 print the state after every layer/gate."""
        self.initialize_state()
        circ = self.circ.build() if not self.circ.built else self.circ
        if debug:
            for layer in circ.layers:
                for op in layer:
                    self.apply_gate(op)
                    if debug == 2:
                        print(self._pretty())
                if debug == 1:
                    print(self._pretty())
        else:
            prog = self.program()
            if progress_bar:
                from tqdm import tqdm
                with tqdm(total=1, desc=f"{prog.n_passes} fused passes") as bar:
                    run_program(self._flat, prog, self._scratch)
                    bar.update(1)
            else:
                run_program(self._flat, prog, self._scratch)
        self.ran = True
        if circ.meas_sites is not None:
            self.measure(circ.meas_sites)      # drawn and discarded, as base_sim.py:49-50
        return self.state

    def get_statevector(self):
        if not self.ran:
            self.run(progress_bar=False)
        return self._flat

    def normalize_state(self):
        lib = _lib.load()
        code = _lib.dtype_code(self.dtype)
        out = torch.empty(1, dtype=torch.float64, device=self.device)
        buf = self._scratch.get(8 * 1024)
        st = _lib.stream_ptr(self.device)
        _lib.check(lib.svh_norm2(_lib.ptr(self._flat), self.num_sites, code, _lib.ptr(out), _lib.ptr(buf),
                                 buf.numel(), st), "svh_norm2")
        _lib.check(lib.svh_scale_by_inv_norm(_lib.ptr(self._flat), self.num_sites, code, _lib.ptr(out), st),
                   "svh_scale_by_inv_norm")

    # ------------------------------------------------------------ measurement
    def measure(self, meas_sites):
        """A measurement bitstring over the sorted sites (statevector_sim.py:54-57)."""
        from .measure import measure_sv
        return measure_sv(self.get_statevector(), meas_sites)

    def measure_and_collapse(self, meas_sites):
        """(bitstring, post-measurement state of the other sites) (statevector_sim.py:59-61)."""
        from .measure import measure_and_collapse_sv
        return measure_and_collapse_sv(self.get_statevector(), meas_sites)

    # ------------------------------------------------------------ observables
    def pauli_expectation(self, pauli_string):
        """Re⟨ψ|P|ψ⟩ for a string over 'ixyz' (statevector_sim.py:75-79,217-234)."""
        if not self.ran:
            self.run(progress_bar=False)
        return float(pauli_sum_expectation(self._flat, self.num_sites, [(1.0, pauli_string)],
                                           self._scratch).item())

    def expectation(self, terms):
        """Σ c·⟨P⟩ over (coeff, pauli string) terms, e.g. Hamiltonian.pauli_terms()."""
        if not self.ran:
            self.run(progress_bar=False)
        return float(pauli_sum_expectation(self._flat, self.num_sites, terms, self._scratch).item())

    def one_local_expectation(self, op, site):
        """Re⟨ψ|O_site|ψ⟩ for a 2x2 O (statevector_sim.py:81-94)."""
        if not self.ran:
            self.run(progress_bar=False)
        right = self._flat.clone()
        from .gates import Gate
        prog = compile_ops(lower_gate(Gate("O", [site], np.asarray(op)), self.num_sites), self.num_sites,
                           self.tile, fuse=False)
        run_program(right, prog, self._scratch)
        return float(inner(self._flat, right, self.num_sites, self._scratch).real.item())

    def local_expectation(self, local_ops):
        """Re⟨ψ|⊗_i O_i|ψ⟩ for one 2x2 per site (statevector_sim.py:63-73)."""
        if not self.ran:
            self.run(progress_bar=False)
        right = self._flat.clone()
        from .gates import Gate
        ops = []
        for i in range(self.num_sites):
            ops.extend(lower_gate(Gate("O", [i], np.asarray(local_ops[i])), self.num_sites))
        prog = compile_ops(ops, self.num_sites, self.tile, fuse=False)
        run_program(right, prog, self._scratch)
        return float(inner(self._flat, right, self.num_sites, self._scratch).real.item())

    def _pretty(self):
        return pretty_print_sv(self._flat.detach().cpu().numpy(), self.local_dim)

    def __str__(self):
        if not self.ran:
            self.run(progress_bar=False)
        return self._pretty()
