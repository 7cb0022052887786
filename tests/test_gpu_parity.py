"""HIP path vs the oracle on the GPU (called through libqaravan_hip.so).

Tolerances: ‖ψ − ψ_ref‖₂ ≤ 1e-12 for complex128 and ≤ 1e-5 for complex64
(plus an n- and depth-aware bound for the 28-qubit c64-vs-c128 check).
"""
import os

import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import CNOT, CZ, Circuit, Gate, H, Reset, X, pauli_Z, random_unitary, string_to_sv
from qaravan_amd import program as P
from qaravan_amd.workloads import brickwork_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C128, C64 = torch.complex128, torch.complex64
TOL = {C128: 1e-12, C64: 1e-5}


def sim_state(circ, gpu, dtype=C64, cfg=None, init=None):
    from qaravan_amd import StatevectorSim
    sim = StatevectorSim(circ, init_state=init, backend="hip", device=gpu, dtype=dtype, tile=cfg)
    sim.run(progress_bar=False)
    return sim, sim.get_statevector().cpu().numpy().astype(np.complex128)


def ghz3():
    return Circuit([H(0), CNOT([1, 0], 3), CNOT([2, 1], 3)], 3)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_ghz_and_observables(gpu, dtype):
    sim, sv = sim_state(ghz3(), gpu, dtype)
    exp = np.zeros(8)
    exp[0] = exp[-1] = 1 / np.sqrt(2)
    assert np.allclose(sv, exp, atol=1e-7)
    assert sim.state.shape == (2, 2, 2)
    assert abs(sim.pauli_expectation('izz') - 1.0) < 1e-7
    assert abs(sim.pauli_expectation('izi')) < 1e-7
    assert abs(sim.one_local_expectation(pauli_Z, 2)) < 1e-7


RESETS = [
    ([H(0), Reset(0, "0")], 1, [1, 0]),
    ([H(0), Reset(0, "1")], 1, [0, 1]),
    ([H(0), X(1), Reset(0, "1")], 2, [0, 0, 0, 1]),
    ([H(0), CNOT([1, 0]), Reset(0, "0")], 2, [1, 0, 0, 0]),
    ([H(0), H(1), Reset([0, 1], "10")], 2, [0, 0, 1, 0]),
    ([H(0), Reset(0, "1"), H(0), Reset(0, "1")], 1, [0, -1]),
    ([H(0), CNOT([1, 0]), Reset(1, "1"), H(1), CNOT([2, 1])], 3,
     (string_to_sv("100") - string_to_sv("111")) / np.sqrt(2)),
]


@pytest.mark.parametrize("case", range(len(RESETS)))
def test_reset_known_answers(gpu, case):
    gates, n, exp = RESETS[case]
    _, sv = sim_state(Circuit(gates, n=n), gpu, C128)
    assert np.allclose(sv, np.asarray(exp, dtype=complex), atol=1e-12)


def test_reset_inside_tile_register(gpu):
    gates = [H(0), CNOT([1, 0]), Reset(1, "1"), H(1), CNOT([2, 1]), H(5), CNOT([9, 5], 10)]
    _, sv = sim_state(Circuit(gates, n=10), gpu, C128)
    assert np.linalg.norm(sv - O.run(gates, 10)) < 1e-12


def test_brickwork_12q_golden(gpu):
    """CPU-sized brickwork (12 qubits, 20 layers, 830 gates) against the fixture."""
    ref = np.load(os.path.join(GOLDEN, "brickwork_12q_20l_state.npz"))["state"]
    circ = brickwork_circuit(12, 20)
    for dtype in (C128, C64):
        _, sv = sim_state(circ, gpu, dtype)
        err = np.linalg.norm(sv - ref)
        print(f"brickwork 12q {dtype}: err={err:.3e}", flush=True)
        assert err <= TOL[dtype]


def random_circuit(n, n_gates, rng, k3=False):
    gates = []
    for _ in range(n_gates):
        r = rng.random()
        if r < 0.35:
            gates.append(Gate("U", [int(rng.integers(n))], random_unitary(2, rng)))
        elif r < 0.6:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(Gate("U", qs, random_unitary(4, rng)))
        elif r < 0.85:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(CNOT(qs, n) if abs(qs[0] - qs[1]) == 1 else CZ(qs))
        else:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(CZ(qs))
    if k3:
        qs = [int(x) for x in rng.choice(n, size=3, replace=False)]
        gates.insert(n_gates // 2, Gate("U3", qs, random_unitary(8, rng)))
    return Circuit(gates, n=n)


CFGS = [P.TileConfig(tb=4, low_bits=6, max_cost=96), P.TileConfig(tb=6, low_bits=4, max_cost=32),
        P.TileConfig(tb=8, low_bits=6, max_cost=96), P.TileConfig(tb=9, low_bits=6, max_cost=1e9),
        P.TileConfig(tb=10, low_bits=5, max_cost=64),
        P.TileConfig(tb=6, low_bits=1, max_cost=48, prefetch=True),
        P.TileConfig(tb=8, low_bits=4, max_cost=1e9, prefetch=True),
        P.TileConfig(tb=9, low_bits=3, max_cost=96, prefetch=True),
        P.TileConfig(tb=8, low_bits=4, max_cost=1e9, specialize=True),
        P.TileConfig(tb=6, low_bits=3, max_cost=48, specialize=True),
        P.TileConfig(tb=8, low_bits=4, max_cost=1e9, specialize=True, io_lane_bits=4, rb=5),
        P.TileConfig(tb=5, low_bits=3, max_cost=48, specialize=True, rb=5),
        P.TileConfig(tb=7, low_bits=4, max_cost=160, specialize=True, io_lane_bits=4, rb=5),
        P.TileConfig(tb=8, low_bits=4, max_cost=160, specialize=True, io_lane_bits=4, rb=5, lds_split=1),
        P.TileConfig(tb=6, low_bits=4, max_cost=1e9, specialize=True, io_lane_bits=4, rb=4, lds_split=1),
        P.TileConfig(tb=7, low_bits=3, max_cost=64, specialize=True, io_lane_bits=3, rb=3, lds_split=1),
        P.TileConfig(tb=8, low_bits=4, max_cost=160, specialize=True, io_lane_bits=4, rb=5, lds_split=2),
        P.TileConfig(tb=9, low_bits=4, max_cost=240, specialize=True, io_lane_bits=4, rb=5, lds_split=2)]


@pytest.mark.parametrize("ci", range(len(CFGS)))
@pytest.mark.parametrize("dtype", [C64, C128])
def test_random_circuits_vs_oracle(gpu, ci, dtype):
    cfg = CFGS[ci]
    if dtype == C128 and (cfg.tb > 9 or (cfg.prefetch and cfg.tb > 8)):
        pytest.skip("c128 tiles stop at 512 threads (256 with prefetch)")
    if dtype == C128 and cfg.tb + cfg.rb > 13:
        pytest.skip("a c128 tile of 2^14 amplitudes needs 256 KiB of LDS unless every pass splits")
    rng = np.random.default_rng(1000 + ci)
    n = max(cfg.tb + cfg.rb, 10) + 2 + (ci % 3)
    circ = random_circuit(n, 120, rng, k3=True)
    print(f"random n={n} cfg={cfg} dtype={dtype}", flush=True)
    _, sv = sim_state(circ, gpu, dtype, cfg)
    ref = O.run(circ.gate_list, n)
    err = np.linalg.norm(sv - ref)
    print(f"  |psi - psi_ref| = {err:.3e}", flush=True)
    assert err <= TOL[dtype]


def test_init_state_forms(gpu):
    rng = np.random.default_rng(3)
    n = 9
    circ = random_circuit(n, 30, rng)
    v = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    v /= np.linalg.norm(v)
    _, sv = sim_state(circ, gpu, C128, init=v)
    assert np.linalg.norm(sv - O.run(circ.gate_list, n, init=v)) < 1e-12
    _, sv = sim_state(circ, gpu, C128, init="101100111")
    assert np.linalg.norm(sv - O.run(circ.gate_list, n, init="101100111")) < 1e-12
    _, sv = sim_state(circ, gpu, C128, init=torch.tensor(v))
    assert np.linalg.norm(sv - O.run(circ.gate_list, n, init=v)) < 1e-12


def test_apply_gate_per_gate_matches_run(gpu):
    from qaravan_amd import StatevectorSim
    rng = np.random.default_rng(9)
    n = 10
    circ = random_circuit(n, 25, rng)
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=C128)
    sim.initialize_state()
    for g in circ.gate_list:
        sim.apply_gate(g)
    a = sim.get_statevector() if sim.ran else sim._flat
    assert np.linalg.norm(a.cpu().numpy() - O.run(circ.gate_list, n)) < 1e-12


def test_expectations_vs_oracle(gpu):
    from qaravan_amd import StatevectorSim
    from qaravan_amd.statevector_sim import pauli_sum_apply
    rng = np.random.default_rng(21)
    n = 11
    circ = random_circuit(n, 60, rng)
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=C128)
    sim.run(progress_bar=False)
    ref = O.run(circ.gate_list, n)
    letters = "ixyz"
    strings = ["".join(letters[int(x)] for x in rng.integers(4, size=n)) for _ in range(12)]
    terms = [(float(rng.standard_normal()), s) for s in strings]
    for c, s in terms:
        assert abs(sim.pauli_expectation(s) - O.pauli_expectation(ref, s)) < 1e-12
    assert abs(sim.expectation(terms) - O.energy(ref, terms)) < 1e-11
    out = torch.empty_like(sim._flat)
    pauli_sum_apply(sim._flat, out, n, terms, sim._scratch)
    assert np.linalg.norm(out.cpu().numpy() - O.apply_pauli_sum(ref, terms)) < 1e-11
    ops = [np.array([[0.3, 0.1 - 0.2j], [0.1 + 0.2j, -0.7]]) for _ in range(n)]
    full = ref.copy()
    for q in range(n):
        full = O.op_action(ops[q], [q], full)
    assert abs(sim.local_expectation(ops) - np.real(np.vdot(ref, full))) < 1e-12


@pytest.mark.parametrize("variant", ["generic", "prefetch", "specialize", "specialize_rb5", "specialize_rb5_half",
                                     "specialize_rb5_quarter"])
def test_brickwork_24q_c64_vs_c128_and_inverse(gpu, variant):
    """Size-independent checks at 24 qubits: c64 agrees with c128, the norm is
    kept, and running the circuit then its inverse returns |0…0⟩."""
    circ = brickwork_circuit(24, 12)
    cfg = P.TileConfig(tb=8, low_bits=4, max_cost=1e9, prefetch=variant == "prefetch",
                       specialize=variant.startswith("specialize"), io_lane_bits=4 if "rb5" in variant else 6,
                       rb=5 if "rb5" in variant else 4,
                       lds_split=1 if variant.endswith("half") else 2 if variant.endswith("quarter") else 0)
    _, a = sim_state(circ, gpu, C128, cfg)
    _, b = sim_state(circ, gpu, C64, cfg)
    assert abs(np.linalg.norm(a) - 1) < 1e-12
    assert np.linalg.norm(a - b) < 1e-5
    inv = Circuit(circ.gate_list + [g.dag() for g in circ.gate_list[::-1]], n=24)
    _, z = sim_state(inv, gpu, C128, cfg)
    assert abs(z[0] - 1) < 1e-11
