"""Monte-Carlo trajectories and Trotter evolution on the GPU.

Mirrors src/qaravan/tensorQ/monte_carlo_sim.py:3-35 and
src/qaravan/applications/trotter.py:11-78:

* ``MonteCarloSim(circ, init_state=None, nm=None)``: ``circuit_sample()``
  builds the circuit and puts one ``nm.sample()`` layer (a list of gates)
  after every circuit layer (monte_carlo_sim.py:14-28);
  ``expectation_sample(obs)`` runs one sampled circuit and returns
  Re⟨ψ|O|ψ⟩ (:30-35).  ``expectation_samples(obs, num_samples)`` runs
  ``num_samples`` trajectories as ONE batched launch sequence (every
  trajectory's state in one flat buffer, pass r of every trajectory in one
  launch) instead of one simulator per sample.
* ``trotter_sv_sim(ham, step_size, max_steps, op)`` (trotter.py:11-23) and
  ``trotter_mc_sim(ham, step_size, max_steps, channel, op, num_samples)``
  (:61-78): the trajectories' states stay on the device between steps.

The noise model ``nm`` is any object whose ``sample()`` returns a list of
gates (as the reference's PauliNoise / PauliLindbladNoise do); sampling runs
on the host with numpy's global RNG, exactly as in the reference, so seeded
runs draw the same circuits.  Observables ``obs`` / ``op`` are Pauli sums
(a list of (coefficient, pauli string) pairs, or an object with
``pauli_terms()``) evaluated by the FP64 expectation kernel; a dense matrix
is accepted for n ≤ 6.  ``magnetization_terms(n)`` is trotter.py:6-9's
operator Σ_i Z_i / sqrt(n·2^n) as Pauli terms.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .batch import BatchedForward
from .program import PASS_TILE, TileConfig, compile_ops, lower_circuit
from .statevector_sim import DeviceScratch, StatevectorSim, inner, pauli_sum_expectation, run_program


def magnetization_terms(n):
    """Σ_i Z_i / sqrt(n·2^n) (trotter.py:6-9) as Pauli terms."""
    c = 1.0 / np.sqrt(n * 2 ** n)
    return [(c, 'i' * i + 'z' + 'i' * (n - 1 - i)) for i in range(n)]


def _as_terms(obs):
    if hasattr(obs, "pauli_terms"):
        return obs.pauli_terms()
    if isinstance(obs, (list, tuple)) and obs and isinstance(obs[0], (list, tuple)):
        return list(obs)
    return None


def expectation_value(flat, n, obs, scratch):
    """Re⟨ψ|O|ψ⟩ as a 0-d float64 device tensor."""
    terms = _as_terms(obs)
    if terms is not None:
        return pauli_sum_expectation(flat, n, terms, scratch)
    mat = np.asarray(obs.toarray() if hasattr(obs, "toarray") else obs, dtype=np.complex128)
    if mat.shape != (2 ** n, 2 ** n):
        raise ValueError(f"observable of shape {mat.shape} does not act on {n} qubits")
    if n > 6:
        raise NotImplementedError("dense observables are supported up to 6 qubits; pass Pauli terms")
    lib = _lib.load()
    right = flat.clone()
    bits = (ctypes.c_int * n)(*[n - 1 - q for q in range(n)])
    m_dev = torch.from_numpy(mat).to(device=flat.device, dtype=flat.dtype)
    _lib.check(lib.svh_apply_dense(_lib.ptr(right), n, _lib.dtype_code(flat.dtype), bits, n, _lib.ptr(m_dev),
                                   _lib.stream_ptr(flat.device)), "svh_apply_dense")
    return inner(flat, right, n, scratch).real


def run_circuits(circuits, states, n, tile=None):
    """Run circuit c on states[c] (a (C, 2^n) device tensor) in place: one
    batched launch sequence when every program is made of tile passes, else
    one program run per circuit."""
    tile = tile or TileConfig(tb=min(8, n - 4), low_bits=4, max_cost=1e9, io_lane_bits=4)
    progs = [compile_ops(lower_circuit(c.build() if not c.built else c, n), n, tile) for c in circuits]
    flat = states.reshape(-1)
    if n >= 8 and all(p.kind == PASS_TILE for q in progs for p in q.passes):
        BatchedForward(progs, n, states.device, states.dtype).run(flat)
        return
    sc = DeviceScratch(states.device)
    for c, p in enumerate(progs):
        run_program(states[c], p, sc)


class MonteCarloSim:
    """Noisy simulation by sampling unitary trajectories (monte_carlo_sim.py:3-35)."""

    def __init__(self, circ, init_state=None, nm=None, device="cuda", dtype=torch.complex64, tile=None):
        _lib.require_gpu()
        self.circ = circ.copy()
        self.num_sites = circ.num_sites
        self.local_dim = circ.local_dim
        self.nm = nm
        self.init_state = init_state
        self.device = torch.device(device)
        self.dtype = dtype
        self.tile = tile
        self.state = None
        self.ran = False
        self._scratch = DeviceScratch(self.device)

    def circuit_sample(self):
        """The circuit with one sampled noise layer after each layer (:14-28)."""
        circ = self.circ.copy()
        if not circ.built:
            circ.build()
        if self.nm is not None:
            noise = [self.nm.sample() for _ in range(len(circ.layers))]
            circ.layers = [circ.layers[i // 2] if i % 2 == 0 else noise[i // 2]
                           for i in range(len(circ.layers) * 2)]
        circ.built = True
        return circ

    def expectation_sample(self, obs):
        """Re⟨ψ|O|ψ⟩ for one sampled trajectory (:30-35)."""
        sim = StatevectorSim(self.circuit_sample(), init_state=self.init_state, backend="hip", device=self.device,
                             dtype=self.dtype, tile=self.tile)
        sim.run(progress_bar=False)
        return float(expectation_value(sim.get_statevector(), self.num_sites, obs, self._scratch).item())

    def _initial_states(self, count):
        n = self.num_sites
        states = torch.zeros((count, 2 ** n), dtype=self.dtype, device=self.device)
        if self.init_state is None:
            states[:, 0] = 1.0
        elif isinstance(self.init_state, str):
            states[:, int(self.init_state, 2)] = 1.0
        else:
            states.copy_(torch.as_tensor(np.asarray(self.init_state)).reshape(1, -1).expand(count, -1))
        return states

    def expectation_samples(self, obs, num_samples):
        """Re⟨ψ_t|O|ψ_t⟩ for ``num_samples`` sampled trajectories (numpy
        array), all trajectories run as one batched launch sequence."""
        circuits = [self.circuit_sample() for _ in range(num_samples)]
        states = self._initial_states(num_samples)
        run_circuits(circuits, states, self.num_sites, self.tile)
        vals = torch.stack([expectation_value(states[t], self.num_sites, obs, self._scratch)
                            for t in range(num_samples)])
        return vals.cpu().numpy()


def trotter_sv_sim(ham, step_size, max_steps, op, device="cuda", dtype=torch.complex64):
    """⟨O⟩ after each first-order Trotter step from |0…0⟩ (trotter.py:11-23)."""
    _lib.require_gpu()
    circ = ham.trotter_circ(step_size, 1)
    n = ham.num_sites
    dev = torch.device(device)
    prog = compile_ops(lower_circuit(circ.build(), n), n,
                       TileConfig(tb=min(8, n - 4), low_bits=4, max_cost=1e9, io_lane_bits=4))
    sc = DeviceScratch(dev)
    state = torch.zeros(2 ** n, dtype=dtype, device=dev)
    state[0] = 1.0
    vals = []
    for _ in range(max_steps):
        run_program(state, prog, sc)
        vals.append(expectation_value(state, n, op, sc))
    return [float(v) for v in torch.stack(vals).cpu().numpy()]


def trotter_mc_sim(ham, step_size, max_steps, channel, op, num_samples=1000, device="cuda",
                   dtype=torch.complex64):
    """Trajectory average of ⟨O⟩ after each noisy Trotter step (trotter.py:61-78):
    every step samples one circuit per trajectory and runs all of them in one
    batched launch sequence on states that stay on the device."""
    circ = ham.trotter_circ(step_size, 1)
    sim = MonteCarloSim(circ, nm=channel, device=device, dtype=dtype)
    n = ham.num_sites
    states = sim._initial_states(num_samples)
    means = []
    for _ in range(max_steps):
        circuits = [sim.circuit_sample() for _ in range(num_samples)]
        run_circuits(circuits, states, n)
        vals = torch.stack([expectation_value(states[t], n, op, sim._scratch) for t in range(num_samples)])
        means.append(float(vals.mean().item()))
    return means
