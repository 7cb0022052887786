"""The oracle's environment optimisation pinned on the CPU."""
import numpy as np

from oracle import sv_oracle as O


def rand_unitary(d, rng):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def rand_sv(n, rng):
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    return v / np.linalg.norm(v)


def brick(n, layers, rng):
    gates = []
    for l in range(layers):
        for i in range(l % 2, n - 1, 2):
            gates.append([rand_unitary(4, rng), [i, i + 1]])
    return gates


def test_environment_trace_identity():
    """Tr(M·env) = ⟨post|(M on the gate's qubits)|pre⟩."""
    rng = np.random.default_rng(1)
    n = 6
    pre, post = rand_sv(n, rng), rand_sv(n, rng)
    for idx in ([2, 4], [4, 2], [0, 5]):
        env = O.partial_overlap(pre, post, skip=idx)
        m = rand_unitary(4, rng)
        # env rows/cols follow the sorted sites; permute M into that order
        if idx != sorted(idx):
            m_sorted = m.reshape(2, 2, 2, 2).transpose(1, 0, 3, 2).reshape(4, 4)
        else:
            m_sorted = m
        assert np.isclose(np.trace(m_sorted @ env), np.vdot(post, O.op_action(m, idx, pre)), atol=1e-13)


def test_update_reaches_nuclear_norm_and_cost_never_increases():
    rng = np.random.default_rng(2)
    n = 6
    target = rand_sv(n, rng)
    gates = brick(n, 4, rng)
    pre, post = rand_sv(n, rng), rand_sv(n, rng)
    env = O.partial_overlap(pre, post, skip=gates[3][1])
    c, _, _ = O.environment_update([list(g) for g in gates], 3, pre, post, target, O.init_state(n), 'increasing')
    assert np.isclose(1 - c, np.linalg.svd(env, compute_uv=False).sum(), atol=1e-12)
    gates, costs = O.environment_state_prep(target, gates, n, sweeps=3)
    assert all(b <= a + 1e-12 for a, b in zip(costs, costs[1:]))
    assert costs[-1] < costs[0]
    sv = O.init_state(n)
    for m, idx in gates:
        sv = O.op_action(m, idx, sv)
    assert np.isclose(costs[-1], 1 - abs(np.vdot(target, sv)), atol=1e-12)


def test_reachable_target_converges():
    """A target made by a circuit of the same shape is reached (cost → 0)."""
    rng = np.random.default_rng(3)
    n = 4
    ref = brick(n, 2, rng)
    target = O.init_state(n)
    for m, idx in ref:
        target = O.op_action(m, idx, target)
    gates, costs = O.environment_state_prep(target, brick(n, 2, rng), n, sweeps=60)
    assert costs[-1] < 1e-6
