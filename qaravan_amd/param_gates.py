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
