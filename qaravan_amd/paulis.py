"""Single-qubit Pauli matrices and the lower-case name map.

Follows src/qaravan/core/paulis.py:5-9 (``pauli_I/X/Y/Z`` and ``pauli_mapping``).
The rotation gates use them as generators and the Pauli-sum observables use the
same letters ('i', 'x', 'y', 'z').
"""
import numpy as np

pauli_I = np.array([[1, 0], [0, 1]], dtype=complex)
pauli_X = np.array([[0, 1], [1, 0]], dtype=complex)
pauli_Y = np.array([[0, -1.j], [1.j, 0]], dtype=complex)
pauli_Z = np.array([[1, 0], [0, -1]], dtype=complex)
pauli_mapping = {'i': pauli_I, 'x': pauli_X, 'y': pauli_Y, 'z': pauli_Z}
