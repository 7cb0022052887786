"""Pauli-sum Hamiltonians on chains, as term lists.

Mirrors the term construction of src/qaravan/core/hamiltonians.py:19-88 for the
``'nn'`` locality on a ``LinearLattice`` (lattices.py:22-41) and ``TFI``
(hamiltonians.py:261-272): terms are ``(coefficient, sites, ops)`` with ops in
'xyz', couplings first, then uniform fields (skipped when |h| <= 1e-10).  The
HIP path evaluates ⟨ψ|H|ψ⟩ and H·ψ from these terms directly; no 2^n×2^n matrix
is formed.
"""
import numpy as np


def chain_pairs(L, periodic=False):
    pairs = [(i, i + 1) for i in range(L - 1)]
    if periodic:
        pairs.append((L - 1, 0))
    return pairs


class Hamiltonian:
    def __init__(self, coupling_types, coupling_strengths, field_types, field_strengths,
                 num_x=None, locality='nn', periodic_x=False, spin_half=False):
        if num_x is None:
            raise ValueError("num_x must be provided (chains only)")
        if locality != 'nn':
            raise NotImplementedError("only nearest-neighbour chains are supported")
        self.num_sites = num_x
        self.spin_half = spin_half
        self.terms = []
        for ops, strength in zip(coupling_types, coupling_strengths):
            for pair in chain_pairs(num_x, periodic_x):
                self.terms.append((strength, pair, [ops[0], ops[1]]))
        for op, strength in zip(field_types, field_strengths):
            if np.isscalar(strength):
                if abs(strength) > 1e-10:
                    for i in range(num_x):
                        self.terms.append((strength, [i], [op]))
            else:
                arr = np.asarray(strength)
                if len(arr) != num_x:
                    raise ValueError(f"Field strength array length {len(arr)} must match "
                                     f"number of sites {num_x}")
                for i, s in enumerate(arr):
                    if abs(s) > 1e-10:
                        self.terms.append((s, [i], [op]))

    def pauli_terms(self):
        """(coefficient, pauli string of length n) pairs, 'i' on idle sites.
        Spin-½ operators scale each term by 2^-(#ops)."""
        out = []
        for c, sites, ops in self.terms:
            s = ['i'] * self.num_sites
            for site, op in zip(sites, ops):
                s[site] = op
            coeff = c * (0.5 ** len(ops) if self.spin_half else 1.0)
            out.append((coeff, ''.join(s)))
        return out


class TFI(Hamiltonian):
    """H = jz Σ Z_i Z_{i+1} + h Σ X_i on an open (or periodic) chain."""

    def __init__(self, num_x, jz=1, h=0, periodic_x=False):
        super().__init__(['zz'], [jz], ['x'], [h], num_x=num_x, locality='nn',
                         periodic_x=periodic_x)
