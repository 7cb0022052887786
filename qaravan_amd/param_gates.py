"""Parametrised rotation gates ``U(θ) = exp(-iθP)`` (no factor ½).

Mirrors src/qaravan/core/param_gates.py:21-116 (``ParamGate``, ``RX``, ``RY``,
``RZ``, ``RXX``, ``RYY``, ``RZZ``, ``CPhase``).  The reference builds the
matrix with ``scipy.linalg.expm`` / ``torch.linalg.matrix_exp``; since every
generator here squares to the identity the exponential has the closed form
``cos θ·I − i sin θ·P``, which is what is evaluated (on the host for a float
angle, as a differentiable torch expression for a tensor angle, and inside the
HIP kernels for angles that live on the device).

Each class also carries ``generator`` (a Pauli word over its indices, or a
diagonal for CPhase) which the adjoint differentiator uses:
``∂U/∂θ = −i·G·U``.
"""
import numpy as np
import torch

from .gates import Gate
from .paulis import pauli_mapping


def _pauli_word(word):
    m = np.array([[1.0 + 0j]])
    for ch in word:
        m = np.kron(m, pauli_mapping[ch])
    return m


class ParamGate(Gate):
    """Gate whose matrix is a function of ``angles`` (param_gates.py:21-49)."""

    generator_word = None   # Pauli word such that U = exp(-iθ P)

    def __init__(self, name, indices, *args):
        super().__init__(name, indices, None)
        self.backend = "torch" if torch.is_tensor(args[0]) else "numpy"
        if isinstance(args[0], np.ndarray) and args[0].ndim == 2:
            self.matrix = args[0]
            self.angles = self.solve_angles()
        else:
            self.angles = args
            self.matrix = self.construct_matrix()

    def __str__(self):
        return f"{self.name} gate on site(s) {self.indices} with angle(s) {self.angles}"

    def update_matrix(self, mat):
        self.matrix = mat

    def solve_angles(self):
        return "unsolved"

    def construct_matrix(self):
        theta = self.angles[0]
        p = _pauli_word(self.generator_word)
        if self.backend == "torch":
            pt = torch.tensor(p, dtype=torch.complex128, device=theta.device)
            eye = torch.eye(p.shape[0], dtype=torch.complex128, device=theta.device)
            th = theta.to(torch.float64)
            return torch.cos(th) * eye - 1j * torch.sin(th) * pt
        th = float(np.real(theta))
        return np.cos(th) * np.eye(p.shape[0], dtype=np.complex128) - 1j * np.sin(th) * p

    @property
    def generator(self):
        return _pauli_word(self.generator_word)

    def shallow_copy(self):
        if isinstance(self.angles, str):
            g = type(self)(self.indices, self.matrix)
        else:
            g = type(self)(self.indices, *self.angles)
            g.matrix = self.matrix
        return g

    def dag(self):
        if isinstance(self.angles, str):
            return Gate.dag(self)
        return type(self)(self.indices, -self.angles[0])


class RX(ParamGate):
    generator_word = "x"

    def __init__(self, indices, *args):
        super().__init__("RX", indices, *args)


class RY(ParamGate):
    generator_word = "y"

    def __init__(self, indices, *args):
        super().__init__("RY", indices, *args)


class RZ(ParamGate):
    generator_word = "z"

    def __init__(self, indices, *args):
        super().__init__("RZ", indices, *args)


class RXX(ParamGate):
    generator_word = "xx"

    def __init__(self, indices, *args):
        super().__init__("RXX", indices, *args)


class RYY(ParamGate):
    generator_word = "yy"

    def __init__(self, indices, *args):
        super().__init__("RYY", indices, *args)


class RZZ(ParamGate):
    generator_word = "zz"

    def __init__(self, indices, *args):
        super().__init__("RZZ", indices, *args)


class CPhase(ParamGate):
    """diag(1, 1, 1, e^{iθ}) (param_gates.py:105-116); generator −|11⟩⟨11|."""

    def __init__(self, indices, *args):
        super().__init__("CPhase", indices, *args)

    def construct_matrix(self):
        theta = self.angles[0]
        if self.backend == "torch":
            th = theta.to(torch.float64)
            d = torch.stack([torch.ones_like(th), torch.ones_like(th), torch.ones_like(th),
                             torch.exp(1j * th)])
            return torch.diag(d.to(torch.complex128))
        return np.diag([1, 1, 1, np.exp(1j * float(theta))]).astype(np.complex128)

    def solve_angles(self):
        return (float(np.angle(self.matrix[3, 3])),)

    @property
    def generator(self):
        return np.diag([0, 0, 0, -1.0]).astype(np.complex128)

    def dag(self):
        if isinstance(self.angles, str):
            return Gate.dag(self)
        return CPhase(self.indices, -self.angles[0])


class U(ParamGate):
    """General single-qubit gate ``U(indices, θ, φ, λ)`` (param_gates.py:118-159):

        [[cos(θ/2),          −e^{iλ} sin(θ/2)],
         [e^{iφ} sin(θ/2),   e^{i(φ+λ)} cos(θ/2)]]

    Angles may be floats or 0-d torch tensors (then the matrix is a
    differentiable complex128 tensor).  Built from a matrix, the angles are
    recovered by ``solve_angles`` after the global phase of the top-left
    entry is divided out of the matrix, as the reference does."""

    def __init__(self, indices, *args):
        super().__init__("U", indices, *args)

    def construct_matrix(self):
        t, p, l = self.angles
        if self.backend == "torch":
            def as_t(a):
                return a.to(torch.float64) if torch.is_tensor(a) else torch.tensor(float(a), dtype=torch.float64)
            t, p, l = as_t(t), as_t(p), as_t(l)
            c, s = torch.cos(t / 2), torch.sin(t / 2)
            e_l, e_p = torch.exp(1j * l), torch.exp(1j * p)
            return torch.stack([torch.stack([c.to(torch.complex128), -e_l * s]),
                                torch.stack([e_p * s, e_p * e_l * c])]).to(torch.complex128)
        t, p, l = (float(np.real(a)) for a in (t, p, l))
        c, s = np.cos(t / 2), np.sin(t / 2)
        return np.array([[c, -np.exp(1j * l) * s], [np.exp(1j * p) * s, np.exp(1j * (p + l)) * c]],
                        dtype=np.complex128)

    def solve_angles(self):
        m = np.asarray(self.matrix, dtype=np.complex128)
        m = m / np.exp(1j * np.angle(m[0, 0]))
        self.matrix = m
        t = 2 * np.arccos(np.clip(m[0, 0].real, -1.0, 1.0))
        if abs(np.sin(t / 2)) > 1e-12:
            l = np.angle(-m[0, 1] / np.sin(t / 2))
            p = np.angle(m[1, 1] / m[0, 0]) - l
        else:
            l = p = np.angle(m[1, 1]) / 2
        return t, p, l

    def decompose(self, basis='ZSX'):
        """Five gates in application order: RZ(λ), SX, RZ(θ+π), SX, RZ(φ+π)
        (param_gates.py:145-159; qaravan's RZ(a) is exp(−i·a·Z))."""
        if basis != 'ZSX':
            raise NotImplementedError(f"{basis} basis decomposition has not been implemented")
        from .gates import SX
        t, p, l = self.solve_angles()
        i = self.indices
        return [RZ(i, l), SX(i), RZ(i, t + np.pi), SX(i), RZ(i, p + np.pi)]

    def dag(self):
        return Gate.dag(self)


def _xyz_center(a, b, c):
    """exp(i(a·XX + b·YY + c·ZZ)).  XX, YY and ZZ commute and square to I, so
    the exponential is the product of cos·I + i·sin·P over the three words;
    written out, it is diagonal in the Bell-type pairs {|00⟩,|11⟩} and
    {|01⟩,|10⟩}:
        ⟨00|·|00⟩ = ⟨11|·|11⟩ = e^{ic} cos(a−b),  ⟨00|·|11⟩ = ⟨11|·|00⟩ = i e^{ic} sin(a−b),
        ⟨01|·|01⟩ = ⟨10|·|10⟩ = e^{−ic} cos(a+b), ⟨01|·|10⟩ = ⟨10|·|01⟩ = i e^{−ic} sin(a+b)."""
    if torch.is_tensor(a) or torch.is_tensor(b) or torch.is_tensor(c):
        def f(x):
            return x.to(torch.float64) if torch.is_tensor(x) else torch.tensor(float(x), dtype=torch.float64)
        a, b, c = f(a), f(b), f(c)
        ep, em = torch.exp(1j * c), torch.exp(-1j * c)
        d0, o0 = ep * torch.cos(a - b), 1j * ep * torch.sin(a - b)
        d1, o1 = em * torch.cos(a + b), 1j * em * torch.sin(a + b)
        z = torch.zeros((), dtype=torch.complex128)
        return torch.stack([torch.stack([d0, z, z, o0]), torch.stack([z, d1, o1, z]),
                            torch.stack([z, o1, d1, z]), torch.stack([o0, z, z, d0])]).to(torch.complex128)
    a, b, c = (float(np.real(x)) for x in (a, b, c))
    ep, em = np.exp(1j * c), np.exp(-1j * c)
    d0, o0 = ep * np.cos(a - b), 1j * ep * np.sin(a - b)
    d1, o1 = em * np.cos(a + b), 1j * em * np.sin(a + b)
    return np.array([[d0, 0, 0, o0], [0, d1, o1, 0], [0, o1, d1, 0], [o0, 0, 0, d0]], dtype=np.complex128)


def kak_unitary(params):
    """Two-qubit gate (U⊗U)·exp(i(p6·XX + p7·YY + p8·ZZ))·(U⊗U) from 15
    angles: params[0:3] and [3:6] the left single-qubit gates, [9:12] and
    [12:15] the right ones (param_gates.py:247-260).  A torch tensor of angles
    gives a differentiable complex128 matrix."""
    if torch.is_tensor(params):
        kron = torch.kron
    else:
        params = np.asarray(params, dtype=np.float64)
        kron = np.kron
    left = kron(U(0, *params[0:3]).matrix, U(0, *params[3:6]).matrix)
    right = kron(U(0, *params[9:12]).matrix, U(0, *params[12:15]).matrix)
    return left @ _xyz_center(params[6], params[7], params[8]) @ right
