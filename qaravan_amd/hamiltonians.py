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
        self.coupling_types = list(coupling_types)
        self.coupling_strengths = list(coupling_strengths)
        self.field_types = list(field_types)
        self.field_strengths = list(field_strengths)
        self.op_grouping = None
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
                    raise ValueError(f"{op} field: {len(arr)} per-site strengths given for a chain of "
                                     f"{num_x} sites")
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


    # ------------------------------------------------------------ Trotter circuits
    def trotter_circ(self, step_size, num_steps, order=1):
        """Circuit of a first- or second-order Trotter product of the groups in
        ``op_grouping`` (hamiltonians.py:112-140)."""
        from .circuits import Circuit
        if self.op_grouping is None:
            raise AssertionError("op_grouping is not defined or is invalid")
        gates = []
        if order == 1:
            layers = [self.layer_from_group(g, step_size, self.num_sites) for g in self.op_grouping]
            for _ in range(num_steps):
                for layer in layers:
                    gates += layer
        elif order == 2:
            l0h = self.layer_from_group(self.op_grouping[0], step_size / 2, self.num_sites)
            l0 = self.layer_from_group(self.op_grouping[0], step_size, self.num_sites)
            l1 = self.layer_from_group(self.op_grouping[1], step_size, self.num_sites)
            gates += l0h
            for _ in range(num_steps - 1):
                gates += l1
                gates += l0
            gates += l1
            gates += l0h
        else:
            raise NotImplementedError("Only first and second order Trotter decomposition is supported")
        return Circuit(gates, self.num_sites)

    def layer_from_group(self, group, time, num_sites):
        """Gates of exp(−i·time·G) for one commuting group G: a field letter
        ('x', 'y', 'z'), a coupling pair ('zz', ...), or 'even'/'odd' bonds
        with every coupling type (hamiltonians.py:142-199)."""
        from scipy.linalg import expm

        from .gates import Gate
        from .paulis import pauli_mapping as P
        half = 0.5 if self.spin_half else 1.0
        gates = []
        if len(group) == 1 and group in self.field_types:
            fs = self.field_strengths[self.field_types.index(group)]
            if not np.isscalar(fs):
                for i, s in enumerate(np.asarray(fs)):
                    if abs(s) > 1e-10:
                        gates.append(Gate('R1', [i], expm(-1j * time * s * P[group])))
                return gates
        if group in ('even', 'odd'):
            gen = sum(c * np.kron(P[t[0]] * half, P[t[1]] * half)
                      for c, t in zip(self.coupling_strengths, self.coupling_types))
        else:
            c = (self.coupling_strengths[self.coupling_types.index(group)] if len(group) == 2
                 else self.field_strengths[self.field_types.index(group)])
            op = np.ones((1, 1), dtype=complex)
            for letter in group:
                op = np.kron(op, P[letter] * half)
            gen = c * op
        mat = expm(-1j * time * np.asarray(gen, dtype=complex))
        if len(group) == 1:
            gates = [Gate('R1', [i], mat) for i in range(num_sites)]
        elif len(group) == 2:
            gates = [Gate('R2', [i, i + 1], mat) for i in range(0, num_sites, 2)]
            gates += [Gate('R2', [i, i + 1], mat) for i in range(1, num_sites - 1, 2)]
        elif group == 'even':
            gates = [Gate('R2', [i, i + 1], mat) for i in range(0, num_sites, 2)]
        elif group == 'odd':
            gates = [Gate('R2', [i, i + 1], mat) for i in range(1, num_sites - 1, 2)]
        else:
            raise NotImplementedError("Only field and nearest-neighbor couplings are supported")
        return gates


class TFI(Hamiltonian):
    """H = jz Σ Z_i Z_{i+1} + h Σ X_i on an open (or periodic) chain
    (Trotter groups: the x field, then the zz bonds)."""

    def __init__(self, num_x, jz=1, h=0, periodic_x=False):
        super().__init__(['zz'], [jz], ['x'], [h], num_x=num_x, locality='nn',
                         periodic_x=periodic_x)
        self.op_grouping = ['x', 'zz']


class Heisenberg(Hamiltonian):
    """H = Σ (jx XX + jy YY + jz ZZ) on bonds + h Σ Z (hamiltonians.py:201-211)."""

    def __init__(self, num_x, jx=1, jy=1, jz=1, h=0):
        super().__init__(['xx', 'yy', 'zz'], [jx, jy, jz], ['z'], [h], num_x=num_x, locality='nn')
        self.op_grouping = ['even', 'odd', 'z']
