"""Batched trajectories, Monte-Carlo sampling and Trotter evolution on the GPU
vs the oracle (through libqaravan_hip.so).  The sampled circuits come from
numpy's global RNG on the host, so re-seeding reproduces them for the oracle.
Tolerances: complex128 ≤ 1e-10, complex64 ≤ 1e-5 on expectation values and
state norms of differences."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import CNOT, Circuit, Gate, random_unitary
from qaravan_amd.batch import BatchedForward
from qaravan_amd.hamiltonians import TFI
from qaravan_amd.monte_carlo import MonteCarloSim, magnetization_terms, trotter_mc_sim, trotter_sv_sim
from qaravan_amd.program import TileConfig, compile_ops, lower_circuit

pytestmark = pytest.mark.gpu
# state tolerances (north star); an energy of a Pauli sum H = Σ c_t P_t differs by at most
# |⟨ψ|H|ψ⟩ − ⟨φ|H|φ⟩| ≤ 2·Σ|c_t|·‖ψ − φ‖₂ for unit vectors, so energies are held to that bound
TOL = {torch.complex128: 1e-12, torch.complex64: 1e-5}


def energy_tol(terms, dtype):
    return 2.0 * sum(abs(c) for c, _ in terms) * TOL[dtype]


class PauliChannel:
    """Independent Pauli noise per qubit: X, Y, Z each with probability p/3."""

    def __init__(self, n, p):
        self.n, self.p = n, p

    def sample(self):
        letters = np.random.choice(list('ixyz'), size=self.n, p=[1 - self.p] + [self.p / 3] * 3)
        return [Gate(ch.upper(), [q], O.PAULI[ch]) for q, ch in enumerate(letters) if ch != 'i']


def random_circuit(n, depth, rng):
    gates = []
    for l in range(depth):
        for q in range(n):
            gates.append(Gate("U", [q], random_unitary(2, rng)))
        for i in range(l % 2, n - 1, 2):
            gates.append(CNOT([i, i + 1], n) if rng.random() < 0.5 else Gate("U2", [i, i + 1], random_unitary(4, rng)))
    return Circuit(gates, n=n)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_batched_forward_matches_oracle(gpu, dtype):
    rng = np.random.default_rng(0)
    n, T = 11, 7
    circs = [random_circuit(n, 2 + t % 4, rng) for t in range(T)]
    progs = [compile_ops(lower_circuit(c.build(), n), n, TileConfig(tb=6, low_bits=4, max_cost=48.0)) for c in circs]
    init = rng.normal(size=(T, 2 ** n)) + 1j * rng.normal(size=(T, 2 ** n))
    init /= np.linalg.norm(init, axis=1, keepdims=True)
    states = torch.tensor(init, dtype=dtype, device=gpu)
    ref_in = states.cpu().numpy().astype(np.complex128)
    BatchedForward(progs, n, gpu, dtype).run(states.reshape(-1))
    for t in range(T):
        ref = O.run(circs[t].gate_list, n, init=ref_in[t])
        err = np.linalg.norm(states[t].cpu().numpy() - ref)
        print(f"  trajectory {t}: |psi - psi_ref| = {err:.3e}", flush=True)
        assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_monte_carlo_samples_match_oracle(gpu, dtype):
    rng = np.random.default_rng(1)
    n = 10
    circ = random_circuit(n, 3, rng)
    obs = TFI(n, jz=-1.0, h=-1.0)
    sim = MonteCarloSim(circ, nm=PauliChannel(n, 0.2), device=gpu, dtype=dtype)
    np.random.seed(42)
    vals = sim.expectation_samples(obs, 12)
    np.random.seed(42)
    samples = [sim.circuit_sample() for _ in range(12)]
    for v, c in zip(vals, samples):
        sv = O.run([g for layer in c.layers for g in layer], n, layered=False)
        assert abs(v - O.energy(sv, obs.pauli_terms())) <= energy_tol(obs.pauli_terms(), dtype)
    np.random.seed(7)
    one = sim.expectation_sample(obs)
    np.random.seed(7)
    c = sim.circuit_sample()
    sv = O.run([g for layer in c.layers for g in layer], n, layered=False)
    assert abs(one - O.energy(sv, obs.pauli_terms())) <= energy_tol(obs.pauli_terms(), dtype)


def test_dense_observable_small_register(gpu):
    rng = np.random.default_rng(2)
    n = 5
    circ = random_circuit(n, 2, rng)
    obs = O.pauli_sum_matrix(magnetization_terms(n), n)
    sim = MonteCarloSim(circ, device=gpu, dtype=torch.complex128)
    v = sim.expectation_samples(obs, 3)
    sv = O.run(circ.gate_list, n)
    assert np.allclose(v, np.vdot(sv, obs @ sv).real, atol=1e-12)


def test_trotter_sv_and_mc(gpu):
    n, dt, steps = 10, 0.1, 4
    ham = TFI(n, jz=1.0, h=0.7)
    op = magnetization_terms(n)
    got = trotter_sv_sim(ham, dt, steps, op, device=gpu, dtype=torch.complex128)
    circ = ham.trotter_circ(dt, 1)
    sv = O.init_state(n)
    ref = []
    for _ in range(steps):
        sv = O.run(circ.gate_list, n, init=sv)
        ref.append(O.energy(sv, op))
    assert np.allclose(got, ref, rtol=0, atol=energy_tol(op, torch.complex128))
    # noisy trajectories: same seed -> same sampled circuits per step
    T = 6
    ch = PauliChannel(n, 0.1)
    np.random.seed(5)
    got = trotter_mc_sim(ham, dt, 3, ch, op, num_samples=T, device=gpu, dtype=torch.complex128)
    np.random.seed(5)
    sim = MonteCarloSim(circ, nm=ch)
    svs = [O.init_state(n) for _ in range(T)]
    ref = []
    for _ in range(3):
        layer = [sim.circuit_sample() for _ in range(T)]
        vals = []
        for t in range(T):
            svs[t] = O.run([g for lay in layer[t].layers for g in lay], n, init=svs[t], layered=False)
            vals.append(O.energy(svs[t], op))
        ref.append(np.mean(vals))
    assert np.allclose(got, ref, rtol=0, atol=energy_tol(op, torch.complex128))
