"""Environment-tensor optimisation and the device op_action vs the oracle
(through libqaravan_hip.so).  Tolerances: complex128 states, ≤ 1e-12 for
single contractions, ≤ 1e-9 on costs and gates after full sweeps (each
update goes through a host SVD of a 4x4 environment)."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd import Circuit, Gate
from qaravan_amd.environments import SweepContext, environment_state_prep, environment_update, sv_environment
from qaravan_amd.statevector_sim import op_action

pytestmark = pytest.mark.gpu


def rand_unitary(d, rng):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def rand_sv(n, rng):
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    return v / np.linalg.norm(v)


def brick_gates(n, layers, rng):
    return [[rand_unitary(4, rng), [i, i + 1]] for l in range(layers) for i in range(l % 2, n - 1, 2)]


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_op_action(gpu, dtype):
    rng = np.random.default_rng(0)
    n = 9
    sv = rand_sv(n, rng)
    t = torch.tensor(sv, dtype=dtype, device=gpu)
    ref_in = t.cpu().numpy().astype(np.complex128)
    for idx in ([3], [7, 2], [0, 8, 4], [1, 2, 3, 5, 6]):
        m = rand_unitary(2 ** len(idx), rng)
        got = op_action(m, idx, t)
        assert got.shape == t.shape and torch.equal(t, torch.tensor(sv, dtype=dtype, device=gpu))
        tol = 1e-12 if dtype == torch.complex128 else 2e-6
        assert np.linalg.norm(got.cpu().numpy() - O.op_action(m, idx, ref_in)) < tol
    r = op_action(rand_unitary(2, rng), [1], t.view([2] * n))
    assert r.shape == (2,) * n


def test_sv_environment(gpu):
    rng = np.random.default_rng(1)
    n = 8
    gl = brick_gates(n, 3, rng)
    circ = Circuit([Gate("U", idx, m) for m, idx in gl], n=n)
    left = rand_sv(n, rng)
    g = 5
    env, mat = sv_environment(circ, left, g, device=gpu)
    sv1 = O.run([Gate("U", idx, m) for m, idx in gl[:g]], n, layered=False)
    sv2 = left.copy()
    for m, idx in reversed(gl[g + 1:]):
        sv2 = O.op_action(m.conj().T, idx, sv2)
    assert np.allclose(env.cpu().numpy(), O.partial_overlap(sv1, sv2, gl[g][1]), atol=1e-12)
    assert np.array_equal(mat, gl[g][0])


def test_environment_state_prep_matches_oracle(gpu):
    rng = np.random.default_rng(2)
    n = 8
    target = rand_sv(n, rng)
    gl = brick_gates(n, 4, rng)
    circ = Circuit([Gate("rand_U", idx, m.copy()) for m, idx in gl], n=n)
    circ, costs = environment_state_prep(target, circ=circ, context=SweepContext(max_iter=2), device=gpu)
    ref_gates, ref_costs = O.environment_state_prep(target, [[m.copy(), idx] for m, idx in gl], n, sweeps=2)
    assert len(costs) == len(ref_costs)
    assert np.allclose(costs, ref_costs, atol=1e-9)
    for g, (m, _) in zip(circ.gate_list, ref_gates):
        assert np.allclose(np.asarray(g.matrix), m, atol=1e-8)
    assert all(b <= a + 1e-10 for a, b in zip(costs, costs[1:]))


def test_environment_update_with_host_arrays(gpu):
    rng = np.random.default_rng(3)
    n = 7
    gl = brick_gates(n, 3, rng)
    circ = Circuit([Gate("U", idx, m.copy()) for m, idx in gl], n=n)
    pre, post, target, init = rand_sv(n, rng), rand_sv(n, rng), rand_sv(n, rng), O.init_state(n)
    for direction, i in (("decreasing", 4), ("increasing", 2), ("decreasing", 1), ("increasing", len(gl) - 2)):
        c, p2, q2 = environment_update(circ, i, pre, post, target, init, direction)
        rc, rp, rq = O.environment_update([[m, idx] for m, idx in gl], i, pre, post, target, init, direction)
        assert abs(c - rc) < 1e-12
        assert np.allclose(p2.cpu().numpy(), rp, atol=1e-12) and np.allclose(q2.cpu().numpy(), rq, atol=1e-12)
