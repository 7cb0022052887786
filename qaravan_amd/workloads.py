"""Seeded benchmark circuits (frozen generators, seeds 20260000 + index).

* ``brickwork_circuit(n, layers, seed)``: every layer applies RX, RY, RZ to
  every qubit (angles uniform in [0, 2π), drawn per layer, qubit, gate) and then
  ``CNOT([i, i+1])`` (target i, control i+1) on even bonds i = 0, 2, … in even
  layers and odd bonds i = 1, 3, … in odd layers.  n=12/20 layers is the CPU
  case (830 gates); n=28/40 layers is the single-GPU case (3900 gates).
* ``haar_block_layers(n, layers, seed)``: layer l tiles qubits
  [o+5j, o+5j+5) ⊂ [0, n) with o = 2l mod 5 by Haar-random 32x32 unitaries
  (QR of a complex Gaussian with the R-phase fix).  n=30/24 layers gives 125
  blocks.
* ``hea_circuit(n, layers, theta)``: hardware-efficient ansatz: per layer
  RY(θ) on every qubit, RZ(θ) on every qubit, then CZ([i, i+1]) for
  i = 0..n-2.  θ index = (2l + k)·n + q (k = 0 RY, 1 RZ).  n=20/16 layers has
  640 parameters and 944 gates.
* ``batch_specs(nc, seed)``: the 512-circuit batch: n ~ U{10..22},
  layers ~ U{4..32}, θ ~ U[−π, π) per circuit (seed 20260004).
* ``tfim_terms(n)``: H = −Σ Z_i Z_{i+1} − Σ X_i on an open chain, i.e.
  qaravan's ``TFI(n, jz=-1, h=-1)`` (core/hamiltonians.py:261-272).
The gates are the package's own Gate / RX / RY / RZ / CNOT / CZ objects, so
the same circuits feed the HIP path and the oracle.
"""
import numpy as np

from .circuits import Circuit
from .gates import CNOT, CZ, Gate, haar_unitary
from .hamiltonians import TFI
from .param_gates import RX, RY, RZ

SEED_BASE = 20260000


def brickwork_circuit(n, layers, seed=SEED_BASE + 1):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2.0 * np.pi, size=(layers, n, 3))
    gates = []
    for l in range(layers):
        for q in range(n):
            gates.append(RX([q], float(ang[l, q, 0])))
            gates.append(RY([q], float(ang[l, q, 1])))
            gates.append(RZ([q], float(ang[l, q, 2])))
        start = 0 if l % 2 == 0 else 1
        for i in range(start, n - 1, 2):
            gates.append(CNOT([i, i + 1], n))
    return Circuit(gates, n=n)


def haar_block_layers(n, layers, seed=SEED_BASE + 2):
    """List of layers; each layer is a list of (first_qubit, 32x32 unitary)."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(layers):
        o = (2 * l) % 5
        layer = []
        j = 0
        while o + 5 * j + 5 <= n:
            layer.append((o + 5 * j, haar_unitary(32, rng)))
            j += 1
        out.append(layer)
    return out


def haar_block_circuit(n, layers, seed=SEED_BASE + 2):
    gates = []
    for layer in haar_block_layers(n, layers, seed):
        for q0, u in layer:
            gates.append(Gate("haar5", list(range(q0, q0 + 5)), u))
    return Circuit(gates, n=n)


def hea_theta(n, layers, seed=SEED_BASE + 3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-np.pi, np.pi, size=2 * layers * n)


def hea_circuit(n, layers, theta):
    """Rotation gates carry ``param_index`` so the compiler can read their angle
    from a device θ vector (differentiable path)."""
    theta = np.asarray(theta) if not hasattr(theta, "detach") else theta
    gates = []
    for l in range(layers):
        for k, cls in enumerate((RY, RZ)):
            for q in range(n):
                p = (2 * l + k) * n + q
                th = theta[p]
                g = cls([q], th if hasattr(th, "detach") else float(th))
                g.param_index = p
                gates.append(g)
        for i in range(n - 1):
            gates.append(CZ([i, i + 1]))
    return Circuit(gates, n=n)


def hea_ops(n, layers):
    """The hardware-efficient ansatz of ``hea_circuit`` as elementary ops
    (device-angle rotations and CZ), built without gate objects: the same op
    list ``lower_circuit(hea_circuit(...), param_mode=True)`` produces."""
    from .program import EOp, PAULI_CODE
    ops = []
    for l in range(layers):
        for k, w in enumerate("yz"):
            for q in range(n):
                ops.append(EOp('rot1', (n - 1 - q,), param=(2 * l + k) * n + q, pauli=PAULI_CODE[w]))
        for i in range(n - 1):
            ops.append(EOp('cz', (n - 1 - i, n - 2 - i)))
    return ops


def tfim_terms(n):
    """(coefficient, Pauli string) list of TFI(n, jz=-1, h=-1)."""
    h = TFI(n, jz=-1, h=-1)
    return h.pauli_terms()


def batch_specs(nc=512, seed=SEED_BASE + 4):
    """(ns, layers, thetas) of the batched-circuit workload."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(10, 23, size=nc)
    ls = rng.integers(4, 33, size=nc)
    thetas = [rng.uniform(-np.pi, np.pi, size=2 * int(n) * int(l)) for n, l in zip(ns, ls)]
    return ns, ls, thetas, rng
