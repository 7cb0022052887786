"""Basis-string helpers (src/qaravan/core/utils.py:6-19, 29-50, 77-90)."""
import numpy as np


def endian_transform(arr):
    """Reverse the qubit order of a statevector or operator (utils.py:6-19)."""
    if arr.ndim == 1:
        n = int(np.log2(len(arr)))
        return arr.reshape([2] * n).transpose(list(range(n))[::-1]).reshape(2 ** n)
    if arr.ndim == 2:
        n = int(np.log2(arr.shape[0]))
        order = list(range(n))[::-1] + list(range(n, 2 * n))[::-1]
        return arr.reshape([2] * 2 * n).transpose(order).reshape(2 ** n, 2 ** n)
    raise ValueError("only works for 1D and 2D arrays")


def base(b, n, length):
    """Base-b digits of n, most significant first, padded to ``length``."""
    digits = []
    while n > 0:
        digits.append(n % b)
        n //= b
    digits = digits[::-1] or [0]
    return [0] * (length - len(digits)) + digits


def string_to_sv(string, local_dim=2):
    """Basis vector for a digit string; first character is the most significant
    digit (utils.py:46-50)."""
    n = len(string)
    sv = np.zeros(local_dim ** n)
    sv[int(string, local_dim)] = 1.0
    return sv


def pretty_print_sv(sv, local_dim=2, threshold=1e-3):
    """``amp|bits⟩ + ...`` for amplitudes above threshold (utils.py:77-90)."""
    sv = np.asarray(sv)
    n = sv.shape[0]
    length = round(np.log(n) / np.log(local_dim))
    terms = []
    for idx in np.where(np.abs(sv) > threshold)[0]:
        bits = ''.join(map(str, base(local_dim, int(idx), length)))
        terms.append(f"{sv[idx]:.4f}|{bits}⟩")
    return " + ".join(terms)
