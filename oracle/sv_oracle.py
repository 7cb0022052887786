"""Numpy complex128 restatement of qaravan's dense statevector path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of the benchmark, never by the product package.

What each function restates (paths relative to the qaravan source tree):

* ``op_action``      src/qaravan/tensorQ/statevector_sim.py:111-140
                     (sort indices, permute the gate tensor to match, contract the
                     gate's input legs with the state's legs, outputs in place).
                     ncon_torch==0.3 (absent here) is replaced by np.tensordot
                     followed by np.moveaxis; the contraction is the same.
* ``construct_layers`` src/qaravan/core/circuits.py:42-63 (greedy ASAP layers).
* ``run``            src/qaravan/core/base_sim.py:30-53 and
                     statevector_sim.py:31-52,102-109 (init |0…0⟩ / bitstring /
                     array, per-gate loop over layers, Reset = projector then
                     renormalise).
* ``circ_to_mat``    src/qaravan/core/circuits.py:177-192 (Kronecker brute force).
* ``embed_full``     independent brute force: the 2^n×2^n matrix of a gate on
                     arbitrary indices built entry by entry from the big-endian
                     convention (qubit q ↔ bit n-1-q).
* ``pauli_expectation`` statevector_sim.py:217-234.
* ``tfi_terms``      src/qaravan/core/hamiltonians.py:19-88,261-272.
* ``partial_overlap`` / ``rdm_from_sv``  statevector_sim.py:155-183 (free legs of
                     the kept sites in ascending site order; rows from sv1,
                     columns from conj(sv2)).
* ``measure_sv``     statevector_sim.py:185-193: probabilities = diag of the rdm
                     over the SORTED sites, outcomes = binary strings of width k,
                     drawn by np.random.choice (legacy global RNG).
* ``choice_indices`` the inverse-CDF steps of numpy's RandomState.choice with
                     p given (cumsum, divide by the last entry, searchsorted
                     side='right' of random_sample draws), made explicit so the
                     GPU sampler can be fed the same uniforms.
* ``environment_update`` / ``environment_state_prep``  tensorQ/environments.py:
                     21-80 on a list of (matrix, indices) gates: env =
                     partial_overlap(pre, post) over the gate's qubits, new
                     gate V·W† from the SVD env = W·S·V†, cost 1 − |Tr(U·env)|,
                     left sweep then right sweep.
* ``u_matrix`` / ``kak_unitary`` / ``two_local_gates`` / ``infidelity``
                     src/qaravan/core/param_gates.py:118-129,247-260,
                     core/circuits.py:206-217 and the loss of
                     examples/state_prep_with_ad.ipynb (cell 1); ``fd4_gradient``
                     the fourth-order central difference used to check gradients.
* ``measure_and_collapse_sv`` statevector_sim.py:195-215, including its order
                     convention: outcome character j is projected onto site
                     meas_sites[j] although the characters were drawn for the
                     sorted sites (the two agree when meas_sites is sorted).

Pinning (each is exercised in tests/test_oracle.py):
* reference known answers: tests/test_core.py:8-41,127-137 (GHZ amplitudes,
  ⟨ZZ⟩, ⟨Z⟩) and tests/test_resets.py:22-122 (seven exact Reset vectors);
* op_action vs embed_full on random gates/indices (n<=8) and circ_to_mat vs
  embed_full on contiguous-index circuits;
* rotation closed forms vs scipy.linalg.expm;
* TFI: ground energy of TFI(4, 1, 1)-type chains vs numpy.linalg.eigh of the
  embedded matrix, ⟨+…+|H|+…+⟩ = h·n, ⟨0…0|H|0…0⟩ = jz·(n-1);
* gradients: adjoint (this file) vs parameter shift (exact for exp(-iθP)) vs
  central finite differences;
* partial_overlap vs the partial trace of the full |sv1⟩⟨sv2| (einsum over a
  brute-force 2^n x 2^n outer product) and closed forms (Bell/GHZ rdms,
  product states, trace = ⟨sv2|sv1⟩);
* choice_indices vs np.random.choice itself under the same seed, and
  measure_sv vs a direct per-outcome probability computation;
* environments: Tr(M·env) = ⟨post|M|pre⟩ for random M (the identity that
  makes V·W† the maximiser), |Tr(U·env)| = Σ singular values for the update,
  and a cost sequence that never increases (block-coordinate ascent);
* kak gates: U(π, 0, π) = X, U(π/2, 0, π) = H, U(0, 0, λ) = diag(1, e^{iλ});
  expm(iπ/4·(XX+YY+ZZ)) = e^{iπ/4}·SWAP; unitarity of kak_unitary at random
  angles; agreement with the package's closed-form centre (an independent
  product formula); fd4_gradient exact on cubics and against analytic
  derivatives of trigonometric test functions; infidelity in [0, 1] and 0 at
  target = ψ(params).
"""
import numpy as np

PAULI = {
    'i': np.eye(2, dtype=np.complex128),
    'x': np.array([[0, 1], [1, 0]], dtype=np.complex128),
    'y': np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    'z': np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


# ---------------------------------------------------------------- gate action
def op_action(op, indices, sv, local_dim=2):
    """ψ'[..i_L..] = Σ_j M[i_L, j_L] ψ[..j_L..] with L = ``indices`` (first index
    is the most significant row bit).  Returns a new array with the input's
    shape (flat in → flat out, rank-n in → rank-n out)."""
    op = np.asarray(op, dtype=np.complex128)
    k = len(indices)
    flat = sv.ndim == 1
    n = int(round(np.log(len(sv)) / np.log(local_dim))) if flat else sv.ndim
    state = np.asarray(sv, dtype=np.complex128).reshape([local_dim] * n)
    t = op.reshape([local_dim] * (2 * k))
    order = sorted(range(k), key=lambda a: indices[a])
    t = np.transpose(t, order + [k + a for a in order])
    locs = [indices[a] for a in order]
    out = np.tensordot(t, state, axes=(list(range(k, 2 * k)), locs))
    out = np.moveaxis(out, list(range(k)), locs)
    return out.reshape(local_dim ** n) if flat else out


def construct_layers(gate_list):
    layers, busy = [], []
    for g in gate_list:
        s = set(g.indices)
        li = len(layers) - 1
        while li >= 0 and not (s & busy[li]):
            li -= 1
        if li < len(layers) - 1:
            layers[li + 1].append(g)
            busy[li + 1] |= s
        else:
            layers.append([g])
            busy.append(set(s))
    return layers


def init_state(n, init=None, local_dim=2):
    if init is None:
        sv = np.zeros(local_dim ** n, dtype=np.complex128)
        sv[0] = 1.0
    elif isinstance(init, str):
        sv = np.zeros(local_dim ** len(init), dtype=np.complex128)
        sv[int(init, local_dim)] = 1.0
    else:
        sv = np.asarray(init, dtype=np.complex128).reshape(-1).copy()
    return sv


def run(gate_list, n, init=None, layered=True):
    """Final flat statevector (complex128)."""
    sv = init_state(n, init)
    seq = [g for layer in construct_layers(gate_list) for g in layer] if layered else gate_list
    for g in seq:
        sv = op_action(np.asarray(g.matrix), g.indices, sv)
        if g.name == "Reset" and getattr(g, "normalize", True):
            sv = sv / np.linalg.norm(sv)
    return sv


# ------------------------------------------------------------- brute forces
def embed_full(mat, indices, n):
    """Full 2^n×2^n matrix of ``mat`` on ``indices`` (big-endian), built entry by
    entry: row/col bit of gate index a is qubit indices[a] = address bit n-1-q."""
    k = len(indices)
    mat = np.asarray(mat, dtype=np.complex128)
    N = 2 ** n
    full = np.zeros((N, N), dtype=np.complex128)
    shifts = [n - 1 - q for q in indices]
    clear = 0
    for s in shifts:
        clear |= 1 << s
    for col in range(N):
        jc = 0
        for a, s in enumerate(shifts):
            jc |= ((col >> s) & 1) << (k - 1 - a)
        base = col & ~clear
        for ir in range(2 ** k):
            amp = mat[ir, jc]
            if amp == 0:
                continue
            row = base
            for a, s in enumerate(shifts):
                row |= ((ir >> (k - 1 - a)) & 1) << s
            full[row, col] += amp
    return full


def circ_to_mat(gate_list, n):
    """Kronecker brute force: per layer pad idle qubits with identity, sort
    gates by first index, Kronecker them, multiply layers right to left."""
    mats = []
    for layer in construct_layers(gate_list):
        used = {q for g in layer for q in g.indices}
        items = [(g.indices[0], np.asarray(g.matrix, dtype=np.complex128)) for g in layer]
        items += [(q, np.eye(2, dtype=np.complex128)) for q in range(n) if q not in used]
        items.sort(key=lambda t: t[0])
        m = items[0][1]
        for _, x in items[1:]:
            m = np.kron(m, x)
        mats.append(m)
    out = mats[0]
    for m in mats[1:]:
        out = m @ out
    return out


# ------------------------------------------------------------- rotations
def rotation(word, theta):
    """exp(-iθP) for a Pauli word P (closed form, P² = I)."""
    p = np.array([[1.0 + 0j]])
    for ch in word:
        p = np.kron(p, PAULI[ch])
    return np.cos(theta) * np.eye(p.shape[0]) - 1j * np.sin(theta) * p


# ------------------------------------------------------------- observables
def pauli_string_matrix(ps):
    m = np.array([[1.0 + 0j]])
    for ch in ps:
        m = np.kron(m, PAULI[ch])
    return m


def apply_pauli_string(sv, ps):
    n = len(ps)
    out = np.asarray(sv, dtype=np.complex128)
    for q, ch in enumerate(ps):
        if ch != 'i':
            out = op_action(PAULI[ch], [q], out)
    return out


def pauli_expectation(sv, ps):
    """Re⟨ψ|P|ψ⟩ (statevector_sim.py:217-234)."""
    return float(np.real(np.vdot(sv, apply_pauli_string(sv, ps))))


def tfi_terms(n, jz=1.0, h=0.0, periodic=False):
    """(coeff, pauli string) list: jz Σ Z_iZ_{i+1} then h Σ X_i."""
    terms = []
    pairs = [(i, i + 1) for i in range(n - 1)] + ([(n - 1, 0)] if periodic else [])
    for a, b in pairs:
        s = ['i'] * n
        s[a] = s[b] = 'z'
        terms.append((jz, ''.join(s)))
    if abs(h) > 1e-10:
        for i in range(n):
            s = ['i'] * n
            s[i] = 'x'
            terms.append((h, ''.join(s)))
    return terms


def apply_pauli_sum(sv, terms):
    out = np.zeros_like(np.asarray(sv, dtype=np.complex128))
    for c, ps in terms:
        out += c * apply_pauli_string(sv, ps)
    return out


def energy(sv, terms):
    return float(np.real(np.vdot(sv, apply_pauli_sum(sv, terms))))


def pauli_sum_matrix(terms, n):
    H = np.zeros((2 ** n, 2 ** n), dtype=np.complex128)
    for c, ps in terms:
        H += c * pauli_string_matrix(ps)
    return H


# ------------------------------------------------------------- gradients
# A parametrised circuit is a list of (kind, qubits, param_index_or_None, matrix)
#   kind = Pauli word ('y', 'z', 'xx', ...) for exp(-iθP) with θ = theta[param]
#   kind = None for a fixed gate given by ``matrix``.
def param_run(spec, theta, n, init=None):
    sv = init_state(n, init)
    for kind, qs, p, mat in spec:
        m = rotation(kind, theta[p]) if kind is not None else mat
        sv = op_action(m, qs, sv)
    return sv


def param_energy(spec, theta, n, terms):
    return energy(param_run(spec, theta, n), terms)


def adjoint_gradient(spec, theta, n, terms):
    """dE/dθ by the adjoint method: ψ and λ = Hψ are walked back through U_k†;
    for a rotation exp(-iθP), dE/dθ_k = 2 Im⟨λ_k|P|ψ_k⟩."""
    psi = param_run(spec, theta, n)
    lam = apply_pauli_sum(psi, terms)
    e = float(np.real(np.vdot(psi, lam)))
    grad = np.zeros(len(theta))
    for kind, qs, p, mat in reversed(spec):
        if kind is not None:
            pm = pauli_string_matrix(kind)
            grad[p] += 2.0 * float(np.imag(np.vdot(lam, op_action(pm, qs, psi))))
            u = rotation(kind, theta[p])
        else:
            u = np.asarray(mat, dtype=np.complex128)
        ud = u.conj().T
        psi = op_action(ud, qs, psi)
        lam = op_action(ud, qs, lam)
    return e, grad


def param_shift_gradient(spec, theta, n, terms, idx=None):
    """Exact for exp(-iθP): E(θ) = a + b cos 2θ + c sin 2θ, so
    dE/dθ = E(θ+π/4) − E(θ−π/4)."""
    idx = range(len(theta)) if idx is None else idx
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        tp = np.array(theta, dtype=np.float64)
        tm = tp.copy()
        tp[i] += np.pi / 4
        tm[i] -= np.pi / 4
        g[j] = param_energy(spec, tp, n, terms) - param_energy(spec, tm, n, terms)
    return g


def fd_gradient(spec, theta, n, terms, eps=1e-3, idx=None):
    """Central differences; for this gate family equals E'(θ)·sin(2ε)/(2ε)."""
    idx = range(len(theta)) if idx is None else idx
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        tp = np.array(theta, dtype=np.float64)
        tm = tp.copy()
        tp[i] += eps
        tm[i] -= eps
        g[j] = (param_energy(spec, tp, n, terms) - param_energy(spec, tm, n, terms)) / (2 * eps)
    return g


# ------------------------------------------------- two-local ansatz (kak gates)
def u_matrix(t, p, l):
    """qaravan's U(θ, φ, λ) (src/qaravan/core/param_gates.py:118-129)."""
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -np.exp(1j * l) * s], [np.exp(1j * p) * s, np.exp(1j * (p + l)) * c]],
                    dtype=np.complex128)


def kak_unitary(params):
    """(U⊗U)·expm(i(p6·XX + p7·YY + p8·ZZ))·(U⊗U) (param_gates.py:247-260),
    the centre by scipy.linalg.expm as in the reference."""
    from scipy.linalg import expm
    p = np.asarray(params, dtype=np.float64)
    words = [np.kron(PAULI[c], PAULI[c]) for c in 'xyz']
    left = np.kron(u_matrix(*p[0:3]), u_matrix(*p[3:6]))
    right = np.kron(u_matrix(*p[9:12]), u_matrix(*p[12:15]))
    return left @ expm(1j * sum(a * w for a, w in zip(p[6:9], words))) @ right


class _G:
    def __init__(self, name, indices, matrix):
        self.name, self.indices, self.matrix = name, list(indices), matrix


def two_local_gates(skeleton, params):
    """The gate list of two_local_circ(skeleton, params) (circuits.py:206-217)."""
    return [_G("rand_U", e, kak_unitary(params[15 * i:15 * (i + 1)])) for i, e in enumerate(skeleton)]


def fidelity(params, skeleton, target, n, init=None):
    """|⟨ψ(params)|target⟩|² for the two-local ansatz."""
    psi = run(two_local_gates(skeleton, params), n, init)
    return abs(np.vdot(psi, target)) ** 2


def infidelity(params, skeleton, target, n, init=None):
    """1 − |⟨ψ(params)|target⟩|², the loss of examples/state_prep_with_ad.ipynb
    (cell 1) and, per input state, applications/compilation.py:73-90.  Its
    gradient is −∇fidelity; differencing the fidelity itself keeps the
    rounding of the constant 1 out of finite differences."""
    return 1.0 - fidelity(params, skeleton, target, n, init)


def fd4_gradient(f, x, idx, h=1e-3):
    """Fourth-order central differences [f(x−2h) − 8f(x−h) + 8f(x+h) − f(x+2h)]/(12h):
    truncation error h⁴·f⁽⁵⁾/30, about 1e-13 relative for the trigonometric
    losses here at h = 1e-3."""
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        vals = []
        for d in (-2, -1, 1, 2):
            xx = np.array(x, dtype=np.float64)
            xx[i] += d * h
            vals.append(f(xx))
        g[j] = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
    return g


# ---------------------------------------------------------------- measurement
def partial_overlap(sv1, sv2, skip=None):
    """M[i, j] = Σ_rest sv1[i, rest] · conj(sv2[j, rest]); i, j run over the
    kept (``skip``) sites in ascending order, big-endian
    (statevector_sim.py:155-180)."""
    n = int(round(np.log2(len(sv1))))
    sites = sorted(skip) if skip is not None else []
    rest = [q for q in range(n) if q not in sites]
    a = np.asarray(sv1, dtype=np.complex128).reshape([2] * n).transpose(sites + rest)
    b = np.asarray(sv2, dtype=np.complex128).reshape([2] * n).transpose(sites + rest)
    k = len(sites)
    a = a.reshape(2 ** k, -1)
    b = b.reshape(2 ** k, -1)
    return a @ b.conj().T


def rdm_from_sv(sv, sites):
    """statevector_sim.py:182-183."""
    return partial_overlap(sv, sv, skip=sites)


def marginal_probabilities(sv, sites):
    """Real diagonal of rdm_from_sv(sv, sites) without forming the matrix
    (sites sorted, big-endian), as measure_sv reads it."""
    n = int(round(np.log2(len(sv))))
    sites = sorted(sites)
    rest = [q for q in range(n) if q not in sites]
    p = np.abs(np.asarray(sv, dtype=np.complex128).reshape([2] * n)) ** 2
    return p.transpose(sites + rest).reshape(2 ** len(sites), -1).sum(axis=1)


def choice_indices(probs, uniforms):
    """numpy RandomState.choice(len(p), p=probs) given its random_sample draws."""
    cdf = np.cumsum(probs)
    cdf /= cdf[-1]
    return np.searchsorted(cdf, uniforms, side='right')


def measure_sv(sv, meas_sites, shots=1, uniforms=None):
    """statevector_sim.py:185-193.  With ``uniforms`` (one per shot) the draw
    is explicit; otherwise they come from np.random.random_sample like
    np.random.choice's."""
    # the reference takes the diagonal of the full rdm; marginal_probabilities
    # is that diagonal (tested) without the 2^k x 2^k matrix
    probs = marginal_probabilities(sv, meas_sites)
    k = len(meas_sites)
    if uniforms is None:
        uniforms = np.random.random_sample() if shots == 1 else np.random.random_sample(shots)
    idx = choice_indices(probs, np.atleast_1d(uniforms))
    out = [np.binary_repr(int(i), width=k) for i in idx]
    return out[0] if shots == 1 else out


def measure_and_collapse_sv(sv, meas_sites, uniforms=None):
    """statevector_sim.py:195-215."""
    sv = np.asarray(sv, dtype=np.complex128)
    n = int(round(np.log2(len(sv))))
    outcome = measure_sv(sv, meas_sites, uniforms=uniforms)
    t = sv.reshape([2] * n)
    sl = [slice(None)] * n
    for j, q in enumerate(meas_sites):
        sl[q] = int(outcome[j])
    collapsed = t[tuple(sl)].reshape(-1).copy()
    return outcome, collapsed / np.linalg.norm(collapsed)


# ---------------------------------------------------------------- environments
def environment_update(gates, gate_idx, pre_state, post_state, target_sv, init_sv, direction='decreasing'):
    """tensorQ/environments.py:21-45 on gates = [[matrix, indices], ...]
    (matrices replaced in place)."""
    indices = list(gates[gate_idx][1])
    env = partial_overlap(pre_state, post_state, skip=indices)
    x, s, yh = np.linalg.svd(env)
    new_mat = yh.conj().T @ x.conj().T
    gates[gate_idx][0] = new_mat
    if direction == 'decreasing':
        if gate_idx == 1:
            pre_state = init_sv
        else:
            m, idx = gates[gate_idx - 1]
            pre_state = op_action(np.asarray(m).conj().T, list(idx), pre_state)
        post_state = op_action(new_mat.conj().T, indices, post_state)
    else:
        pre_state = op_action(new_mat, indices, pre_state)
        if gate_idx == len(gates) - 2:
            post_state = target_sv
        else:
            m, idx = gates[gate_idx + 1]
            post_state = op_action(np.asarray(m), list(idx), post_state)
    return 1 - np.abs(np.trace(new_mat @ env)), pre_state, post_state


def environment_state_prep(target_sv, gates, n, init_sv=None, sweeps=1):
    """tensorQ/environments.py:47-80 for a fixed number of sweeps."""
    init = init_state(n, init_sv)
    sv = init.copy()
    for m, idx in gates:
        sv = op_action(np.asarray(m), list(idx), sv)
    cost_list = [1 - np.abs(np.vdot(target_sv, sv))]
    m, idx = gates[-1]
    pre_state = op_action(np.asarray(m).conj().T, list(idx), sv)
    post_state = np.asarray(target_sv, dtype=np.complex128)
    for _ in range(sweeps):
        for i in reversed(range(1, len(gates))):
            c, pre_state, post_state = environment_update(gates, i, pre_state, post_state, target_sv, init,
                                                          'decreasing')
            cost_list.append(c)
        for i in range(len(gates) - 1):
            c, pre_state, post_state = environment_update(gates, i, pre_state, post_state, target_sv, init,
                                                          'increasing')
            cost_list.append(c)
    return gates, cost_list
