"""Measurement, sampling, collapse and partial overlaps on the GPU vs the
oracle (through libqaravan_hip.so).

Tolerances: probabilities and overlaps are FP64 sums of the same amplitudes
the oracle sees (a complex64 state is compared with the oracle run on that
state cast to complex128), so ≤ 1e-12 absolute; sampled outcomes must be
IDENTICAL to the oracle's (numpy's) under the same global seed for
complex128 states.
"""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import measure as M

pytestmark = pytest.mark.gpu
C128, C64 = torch.complex128, torch.complex64


def rand_state(n, seed, dtype, gpu):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    t = torch.tensor(v, dtype=dtype, device=gpu)
    return t, t.cpu().numpy().astype(np.complex128)


SITE_SETS = [(10, []), (10, [3]), (10, [9, 0]), (11, [2, 7, 4]), (12, list(range(12))),
             (12, [11, 5, 0, 6, 1]), (16, list(range(2, 15))), (18, [17]), (14, list(range(13, 0, -1)))]


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,sites", SITE_SETS)
def test_marginal_probabilities(gpu, dtype, n, sites):
    t, ref_sv = rand_state(n, 1000 + n + len(sites), dtype, gpu)
    got = M.marginal_probabilities(t, sites).cpu().numpy()
    assert np.allclose(got, O.marginal_probabilities(ref_sv, sites), rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,skip", [(8, []), (8, [0]), (9, [8]), (10, [4, 1]), (10, [2, 7, 9]),
                                    (12, [0, 5, 6, 11]), (12, [10, 3, 1, 7, 5]), (13, [0, 2, 4, 6, 8, 10]),
                                    (16, [7, 8])])
def test_partial_overlap(gpu, dtype, n, skip):
    a, a_ref = rand_state(n, 2000 + n, dtype, gpu)
    b, b_ref = rand_state(n, 3000 + n, dtype, gpu)
    got = M.partial_overlap(a, b, skip=skip).cpu().numpy()
    assert got.shape == (2 ** len(skip),) * 2
    assert np.allclose(got, O.partial_overlap(a_ref, b_ref, skip), rtol=0, atol=1e-12)
    rdm = M.rdm_from_sv(a, skip).cpu().numpy()
    assert np.allclose(rdm, O.rdm_from_sv(a_ref, skip), rtol=0, atol=1e-12)


def test_rdm_properties_at_22_qubits(gpu):
    from qaravan_amd import StatevectorSim
    from qaravan_amd.workloads import brickwork_circuit
    sim = StatevectorSim(brickwork_circuit(22, 6), backend="hip", device=gpu, dtype=C64)
    sv = sim.get_statevector()
    for sites in ([0, 21], [10, 11, 12], [3, 9, 14, 20]):
        r = M.rdm_from_sv(sv, sites).cpu().numpy()
        assert np.allclose(r, r.conj().T, atol=1e-12)
        assert abs(np.trace(r).real - 1.0) < 1e-5
        assert np.linalg.eigvalsh(r).min() > -1e-6
        p = M.marginal_probabilities(sv, sites).cpu().numpy()
        assert np.allclose(p, np.real(np.diag(r)), atol=1e-12)


@pytest.mark.parametrize("n,sites,shots", [(10, [0, 3], 1), (10, [9, 2, 5], 500), (12, list(range(12)), 2000),
                                           (20, list(range(20)), 10000), (16, [15], 300), (8, [], 5)])
def test_sampling_matches_numpy_choice(gpu, n, sites, shots):
    t, ref_sv = rand_state(n, 4000 + n, C128, gpu)
    for seed in (0, 7):
        np.random.seed(seed)
        got = M.measure_sv(t, sites, shots=shots)
        np.random.seed(seed)
        ref = O.measure_sv(ref_sv, sites, shots=shots)
        assert got == ref
        # the global RNG advanced by the same number of draws
        np.random.seed(seed)
        M.measure_sv(t, sites, shots=shots)
        a = np.random.random_sample()
        np.random.seed(seed)
        O.measure_sv(ref_sv, sites, shots=shots)
        assert a == np.random.random_sample()


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,sites", [(9, [2, 6]), (9, [6, 2]), (10, [0]), (8, list(range(8))), (11, [10, 0, 5])])
def test_measure_and_collapse(gpu, dtype, n, sites):
    t, ref_sv = rand_state(n, 5000 + n, dtype, gpu)
    for seed in (1, 2, 3):
        np.random.seed(seed)
        out, col = M.measure_and_collapse_sv(t, sites)
        np.random.seed(seed)
        ref_out, ref_col = O.measure_and_collapse_sv(ref_sv, sites)
        assert out == ref_out
        tol = 1e-12 if dtype == C128 else 1e-6
        assert col.dtype == dtype and col.numel() == 2 ** (n - len(sites))
        assert np.allclose(col.cpu().numpy(), ref_col, atol=tol)


def test_simulator_measure_and_run_with_meas_sites(gpu):
    from qaravan_amd import CNOT, Circuit, H, StatevectorSim
    n = 4
    circ = Circuit([H(0), CNOT([1, 0], n), CNOT([2, 1], n), CNOT([3, 2], n)], n)
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=C128)
    sim.run(progress_bar=False)
    np.random.seed(3)
    outs = {sim.measure([0, 1, 2, 3]) for _ in range(40)}
    assert outs == {'0000', '1111'}
    out, col = sim.measure_and_collapse([0, 2])
    assert out in ('00', '11')
    exp = np.zeros(4, dtype=complex)
    exp[0 if out == '00' else 3] = 1.0
    assert np.allclose(col.cpu().numpy(), exp)
    circ2 = Circuit([H(0), CNOT([1, 0], n)], n, meas_sites=[0, 1])
    sim2 = StatevectorSim(circ2, backend="hip", device=gpu, dtype=C64)
    st = sim2.run(progress_bar=False)
    assert st.shape == (2,) * n and sim2.ran


def test_pauli_expectation_sv(gpu):
    t, ref_sv = rand_state(9, 77, C128, gpu)
    for ps in ("ixyzizyxi", "zzzzzzzzz", "iiiiiiiii", "xiiiiiiiy"):
        assert abs(M.pauli_expectation_sv(t, ps) - O.pauli_expectation(ref_sv, ps)) < 1e-12
    with pytest.raises(AssertionError):
        M.pauli_expectation_sv(t, "xx")


def test_measurement_errors(gpu):
    t, _ = rand_state(8, 9, C128, gpu)
    with pytest.raises(ValueError):
        M.measure_sv(t, [1, 1])
    with pytest.raises(ValueError):
        M.measure_sv(t, [8])
    with pytest.raises(ValueError):
        M.partial_overlap(t, t, skip=list(range(7)))
    with pytest.raises(ValueError):
        M.measure_sv(t * 2.0, [0])          # probabilities do not sum to 1
