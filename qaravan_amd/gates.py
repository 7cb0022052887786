"""Gate records for the statevector path.

Mirrors the reference data model in src/qaravan/core/gates.py:
``Gate`` (32-55), ``CNOTGate`` (137-170) with ``indices = (target, control)``,
``Reset`` (172-187) and the fixed qubit gates (193-236).  Matrices are numpy
arrays whose row/column bits follow ``gate.indices`` order, first index most
significant, exactly as the reference's ``op_action`` reads them
(src/qaravan/tensorQ/statevector_sim.py:124-140).
"""
from functools import reduce

import numpy as np

from .paulis import pauli_Y, pauli_Z


class Gate:
    """A named operator on ``indices`` (gates.py:32-55)."""

    def __init__(self, name, indices, matrix, time=None):
        self.name = name
        self.indices = [indices] if isinstance(indices, (int, np.integer)) else list(indices)
        self.indices = [int(i) for i in self.indices]
        self.span = len(self.indices)
        self.matrix = matrix
        self.time = time

    def __str__(self):
        return f"{self.name} gate on site(s) {self.indices}"

    def get_local_dim(self):
        # gates.py:46-47: d = D ** (1/span); rounded so it is exact for span <= 8
        return int(round(self.matrix.shape[0] ** (1.0 / self.span)))

    def dag(self):
        g = self.shallow_copy()
        m = self.matrix
        g.matrix = m.conj().T
        return g

    def shallow_copy(self):
        return Gate(self.name, self.indices, self.matrix, time=self.time)


class CNOTGate(Gate):
    """Controlled-X with ``indices = (target, control)`` (gates.py:137-170).

    The stored indices are sorted ascending; when the target is the larger
    index the qubit order of the 4x4 matrix is swapped (``bottom_heavy``), so
    the matrix always reads in ``self.indices`` order.  Non-adjacent pairs use
    the periodic-wrap rule of gates.py:143-152 with chain length ``n``.
    """

    def __init__(self, indices, matrix, n):
        d = int(round(np.sqrt(matrix.shape[0])))
        indices = [int(i) for i in indices]
        if abs(indices[0] - indices[1]) > 1:
            t0, t1 = (indices[0] + 1) % n, (indices[1] + 1) % n
            keep = t0 < t1
        else:
            keep = indices[0] < indices[1]
        if keep:
            mat = matrix
            bottom_heavy = False
        else:
            indices = indices[::-1]
            mat = matrix.reshape(d, d, d, d).transpose(1, 0, 3, 2).reshape(d * d, d * d)
            bottom_heavy = True
        super().__init__(f"CNOT{d}", indices, mat)
        self.bottom_heavy = bottom_heavy
        self._n = n

    def shallow_copy(self):
        g = Gate.shallow_copy(self)
        g.__class__ = CNOTGate
        g.bottom_heavy = self.bottom_heavy
        g._n = self._n
        return g

    def __str__(self):
        title = "bottom heavy" if self.bottom_heavy else "top heavy"
        return f"{title} {self.name} gate on site(s) {self.indices}"


class Reset(Gate):
    """Projector onto ``reset_to`` on ``indices``, then renormalisation
    (gates.py:172-187, base_sim.py:25-28)."""

    def __init__(self, indices, reset_to: str, normalize=True, time=None):
        p0 = np.array([[1, 0], [0, 0]])
        p1 = np.array([[0, 0], [0, 1]])
        mat = reduce(np.kron, [p0 if b == '0' else p1 for b in reset_to])
        super().__init__("Reset", indices, mat, time=time)
        self.reset_to = reset_to
        self.normalize = normalize

    def __str__(self):
        return f"Reset on site(s) {self.indices} to state |{self.reset_to}⟩"

    def shallow_copy(self):
        return Reset(self.indices, self.reset_to, self.normalize, self.time)


# fixed qubit gates (gates.py:193-236)
def SX(indices):
    return Gate("SX", indices, (1 / np.sqrt(2)) * np.array([[1, -1j], [-1j, 1]]))


def X(indices):
    return Gate("X", indices, np.array([[0, 1], [1, 0]]))


def Y(indices):
    return Gate("Y", indices, pauli_Y)


def Z(indices):
    return Gate("Z", indices, pauli_Z)


def H(indices):
    return Gate("H", indices, np.array([[1, 1], [1, -1]]) / np.sqrt(2))


def SDG(indices):
    return Gate("SDG", indices, np.array([[1, 0], [0, -1j]]))


_CNOT_MAT = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]])


def CNOT(indices, n=1000):
    """indices = (target, control)"""
    return CNOTGate(indices, _CNOT_MAT, n=n)


def CZ(indices):
    return Gate("CZ", indices, np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]))


def SWAP(indices):
    return Gate("SWAP", indices, np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]))


def iSWAP(indices):
    return Gate("iSWAP", indices, np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]))


def proj_up(indices):
    return Gate("P0", indices, np.array([[1, 0], [0, 0]]))


def proj_down(indices):
    return Gate("P1", indices, np.array([[0, 0], [0, 1]]))


def random_unitary(size, rng=None):
    """Unitary from the eigenvectors of a·a† with uniform entries
    (gates.py:289-293).  Not Haar distributed; see ``haar_unitary``."""
    r = np.random if rng is None else rng
    a = r.random((size, size)) + 1.j * r.random((size, size))
    _, u = np.linalg.eigh(a @ a.conj().T)
    return u


def haar_unitary(size, rng):
    """Haar-random unitary by QR of a complex Gaussian matrix with the phase
    fix Q <- Q diag(R_ii/|R_ii|) (Mezzadri, Notices AMS 54 (2007) 592)."""
    z = (rng.standard_normal((size, size)) + 1j * rng.standard_normal((size, size))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))
