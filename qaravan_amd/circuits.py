"""Circuit container and greedy ASAP layering.

Mirrors src/qaravan/core/circuits.py: ``Circuit`` (9-24), ``decompose`` (31-40),
``construct_layers`` (42-63), ``build`` (86-95), ``dag`` (112-114), slicing/len/add/mul (126-153),
``copy`` (163-170), ``compose_circuits`` (172-175) and the layer builders
``two_local_circ`` / ``rx_layer`` / ``ry_layer`` / ``rz_layer`` (206-232).
Noise insertion (``add_noise``) is not part of the statevector path, which is
always noiseless.
"""
import numpy as np

from .gates import Gate, random_unitary
from .param_gates import RX, RY, RZ, kak_unitary


class Circuit:
    def __init__(self, gate_list, n=None, local_dim=2, meas_sites=None, requires_decomp=False):
        self.gate_list = gate_list
        if len(gate_list) != 0:
            top = max(max(g.indices) for g in gate_list)
            self.num_sites = int(top + 1) if n is None else n
            self.local_dim = gate_list[0].get_local_dim()
        else:
            self.num_sites = n
            self.local_dim = local_dim
        self.layers = None
        self.built = False
        self.meas_sites = meas_sites
        self.requires_decomp = requires_decomp

    def update_gate(self, gate_idx, new_array):
        self.gate_list[gate_idx].matrix = new_array
        if self.layers is not None:
            self.construct_layers()

    def decompose(self, basis='ZSX'):
        """Replace every ``U`` gate that has angles by its five-gate ZSX
        decomposition (circuits.py:31-40); other gates are kept."""
        out = []
        for g in self.gate_list:
            if g.name == "U" and getattr(g, "angles", None) is not None:
                out.extend(g.decompose(basis))
            else:
                out.append(g)
        self.gate_list = out

    def construct_layers(self):
        """Greedy ASAP layering: a gate goes into the layer after the newest layer
        that shares a qubit with it (circuits.py:42-63).  Gates that overlap keep
        their order; disjoint gates may move earlier."""
        layers = []
        busy = []                      # qubits used by each layer
        for gate in self.gate_list:
            g = set(gate.indices)
            li = len(layers) - 1
            while li >= 0 and not (g & busy[li]):
                li -= 1
            if li < len(layers) - 1:
                layers[li + 1].append(gate)
                busy[li + 1] |= g
            else:
                layers.append([gate])
                busy.append(set(g))
        self.layers = layers

    def build(self, nm=None):
        if nm is not None:
            raise NotImplementedError("the statevector path is noiseless (nm must be None)")
        if not self.built:
            if self.requires_decomp:
                self.decompose()
            self.construct_layers()
            self.built = True
        return self

    def add_meas(self, meas_sites):
        self.meas_sites = meas_sites

    def dag(self):
        return Circuit([g.dag() for g in self.gate_list[::-1]], n=self.num_sites,
                       local_dim=self.local_dim)

    def __str__(self):
        if self.layers is not None:
            return " \n\n".join(" \n".join(str(g) for g in layer) for layer in self.layers)
        return " \n".join(str(g) for g in self.gate_list)

    def __getitem__(self, key):
        if isinstance(key, int):
            return self.gate_list[key]
        if isinstance(key, slice):
            return Circuit(self.gate_list[key], n=self.num_sites, local_dim=self.local_dim)
        raise TypeError("Invalid index type.")

    def __len__(self):
        return len(self.gate_list)

    def __add__(self, other):
        return compose_circuits([self, other])

    def __mul__(self, other):
        if isinstance(other, int) and other >= 0:
            if other == 0:
                return Circuit([], n=self.num_sites, local_dim=self.local_dim)
            if other == 1:
                return self.copy()
            return compose_circuits([self] * other)
        raise TypeError("Circuit multiplication only supports non-negative integers")

    __rmul__ = __mul__

    def copy(self):
        new = Circuit([g.shallow_copy() for g in self.gate_list], n=self.num_sites,
                      local_dim=self.local_dim, meas_sites=self.meas_sites)
        if self.layers is not None:
            new.layers = [[g.shallow_copy() for g in layer] for layer in self.layers]
        new.built = self.built
        return new


def compose_circuits(circ_list):
    gates = [g for c in circ_list for g in c.gate_list]
    return Circuit(gates, n=int(np.max([c.num_sites for c in circ_list])))


def two_local_circ(skeleton, params=None, mag=None, rng=None):
    """One two-qubit gate per skeleton edge (circuits.py:206-217): with
    ``params``, edge i gets ``kak_unitary(params[15i:15i+15])`` (a torch
    tensor of params gives differentiable gate matrices); with only ``mag``,
    ``kak_unitary`` of 15 uniform angles in [0, mag); with neither, a random
    unitary (``random_unitary(4)``).  ``rng`` (a numpy Generator) replaces
    numpy's global RNG for the random draws."""
    gates = []
    for i, e in enumerate(skeleton):
        if params is not None:
            mat = kak_unitary(params[15 * i:15 * (i + 1)])
        elif mag is not None:
            r = np.random if rng is None else rng
            mat = kak_unitary(r.random(15) * mag)
        else:
            mat = random_unitary(4, rng)
        gates.append(Gate("rand_U", list(e), mat))
    return Circuit(gates)


def rx_layer(n, params=None):
    params = np.random.rand(n) * np.pi if params is None else params
    return Circuit([RX([i], params[i]) for i in range(n)])


def ry_layer(n, params=None):
    params = np.random.rand(n) * np.pi if params is None else params
    return Circuit([RY([i], params[i]) for i in range(n)])


def rz_layer(n, params=None):
    params = np.random.rand(n) * np.pi if params is None else params
    return Circuit([RZ([i], params[i]) for i in range(n)])
