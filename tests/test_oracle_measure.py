"""The oracle's measurement functions pinned on the CPU (no GPU): brute-force
partial traces, closed forms, and numpy's own np.random.choice."""
import numpy as np
import pytest

from oracle import sv_oracle as O


def rand_sv(n, rng):
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    return v / np.linalg.norm(v)


def brute_partial(sv1, sv2, skip):
    """Partial trace of the full outer product |sv1><sv2| (2^n x 2^n), traced
    over the non-kept sites with einsum; kept legs in ascending site order."""
    n = int(np.log2(len(sv1)))
    rho = np.outer(sv1, sv2.conj()).reshape([2] * (2 * n))
    keep = sorted(skip)
    letters = 'abcdefghijklmnopqrstuvwxyz'
    row = list(letters[:n])
    col = list(letters[n:2 * n])
    for q in range(n):
        if q not in keep:
            col[q] = row[q]          # trace this site
    out = ''.join(row[q] for q in keep) + ''.join(col[q] for q in keep)
    m = np.einsum(''.join(row) + ''.join(col) + '->' + out, rho)
    k = len(keep)
    return m.reshape(2 ** k, 2 ** k)


@pytest.mark.parametrize("skip", [[0], [3], [1, 4], [4, 1], [0, 2, 5], [5, 3, 1, 0], []])
def test_partial_overlap_vs_partial_trace(skip):
    rng = np.random.default_rng(7)
    a, b = rand_sv(6, rng), rand_sv(6, rng)
    got = O.partial_overlap(a, b, skip)
    assert np.allclose(got, brute_partial(a, b, skip), atol=1e-13)
    # trace over the kept sites gives <b|a>
    assert np.isclose(np.trace(got), np.vdot(b, a), atol=1e-13)


def test_rdm_closed_forms():
    bell = np.zeros(4, dtype=complex)
    bell[0] = bell[3] = 1 / np.sqrt(2)
    assert np.allclose(O.rdm_from_sv(bell, [0]), np.eye(2) / 2)
    assert np.allclose(O.rdm_from_sv(bell, [1]), np.eye(2) / 2)
    # product state |φ0>⊗|φ1>⊗|φ2>: rdm of site 1 is |φ1><φ1|
    rng = np.random.default_rng(3)
    phis = [rand_sv(1, rng) for _ in range(3)]
    sv = np.kron(np.kron(phis[0], phis[1]), phis[2])
    assert np.allclose(O.rdm_from_sv(sv, [1]), np.outer(phis[1], phis[1].conj()))
    assert np.allclose(O.rdm_from_sv(sv, [0, 2]),
                       np.outer(np.kron(phis[0], phis[2]), np.kron(phis[0], phis[2]).conj()))
    r = O.rdm_from_sv(rand_sv(5, rng), [1, 3])
    assert np.allclose(r, r.conj().T) and np.isclose(np.trace(r).real, 1.0)
    assert np.linalg.eigvalsh(r).min() > -1e-14


def test_marginal_is_rdm_diagonal():
    rng = np.random.default_rng(11)
    sv = rand_sv(7, rng)
    for sites in ([2], [6, 0], [1, 3, 4], list(range(7))):
        assert np.allclose(O.marginal_probabilities(sv, sites), np.real(np.diag(O.rdm_from_sv(sv, sites))))
    # per-outcome brute force: p(m) = Σ |ψ_i|² over i whose kept bits spell m
    sites = [5, 1]
    p = np.zeros(4)
    for i, a in enumerate(sv):
        bits = format(i, '07b')
        p[int(bits[1] + bits[5], 2)] += abs(a) ** 2
    assert np.allclose(O.marginal_probabilities(sv, sites), p)


def test_choice_indices_is_numpys_choice():
    rng = np.random.default_rng(5)
    p = rng.random(32)
    p /= p.sum()
    for seed in range(5):
        np.random.seed(seed)
        ref = np.random.choice(32, p=p, size=200)
        np.random.seed(seed)
        got = O.choice_indices(p, np.random.random_sample(200))
        assert np.array_equal(ref, got)
    np.random.seed(9)
    ref1 = np.random.choice(32, p=p)
    np.random.seed(9)
    assert O.choice_indices(p, np.atleast_1d(np.random.random_sample()))[0] == ref1


def test_measure_sv_matches_reference_draw():
    """measure_sv as the reference writes it: np.random.choice over binary
    strings with p = diag(rdm), same global seed."""
    rng = np.random.default_rng(17)
    sv = rand_sv(6, rng)
    for sites, shots in (([0, 3], 1), ([4, 1, 2], 50), ([5], 7)):
        probs = np.real(np.diag(O.rdm_from_sv(sv, sites)))
        outcomes = [np.binary_repr(i, width=len(sites)) for i in range(len(probs))]
        np.random.seed(123)
        ref = str(np.random.choice(outcomes, p=probs)) if shots == 1 else \
            [str(o) for o in np.random.choice(outcomes, p=probs, size=shots)]
        np.random.seed(123)
        assert O.measure_sv(sv, sites, shots=shots) == ref


def test_ghz_measurement_and_collapse():
    n = 5
    ghz = np.zeros(2 ** n, dtype=complex)
    ghz[0] = ghz[-1] = 1 / np.sqrt(2)
    outs = O.measure_sv(ghz, list(range(n)), shots=64, uniforms=np.linspace(0.001, 0.999, 64))
    assert set(outs) == {'0' * n, '1' * n}
    out, col = O.measure_and_collapse_sv(ghz, [1, 3], uniforms=np.array([0.9]))
    assert out == '11'
    expect = np.zeros(8, dtype=complex)
    expect[-1] = 1.0
    assert np.allclose(col, expect)


def test_collapse_is_projection():
    """Sorted sites: collapsed state = (⟨m| on the sites) ψ / norm, with the
    projection built from embed_full of |m⟩⟨m|."""
    rng = np.random.default_rng(23)
    n = 5
    sv = rand_sv(n, rng)
    sites = [1, 4]
    out, col = O.measure_and_collapse_sv(sv, sites, uniforms=np.array([0.37]))
    m = int(out, 2)
    proj = np.zeros((4, 4))
    proj[m, m] = 1.0
    full = O.embed_full(proj, sites, n) @ sv
    kept = full.reshape([2] * n)[(slice(None), int(out[0]), slice(None), slice(None), int(out[1]))].reshape(-1)
    assert np.allclose(col, kept / np.linalg.norm(kept))
