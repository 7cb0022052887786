"""MI355X (gfx950) statevector backend with qaravan's Circuit / StatevectorSim API.

Gate application, observables and the adjoint gradient run in hand-written HIP
kernels (libqaravan_hip.so, C ABI in include/qaravan_hip.h); PyTorch supplies
device memory, streams and autograd only.
"""
from .circuits import Circuit, compose_circuits, rx_layer, ry_layer, rz_layer, two_local_circ
from .gates import (CNOT, CZ, H, SDG, SWAP, SX, CNOTGate, Gate, Reset, X, Y, Z, haar_unitary, iSWAP,
                    proj_down, proj_up, random_unitary)
from .hamiltonians import TFI, Hamiltonian
from .param_gates import RX, RXX, RY, RYY, RZ, RZZ, CPhase, ParamGate
from .paulis import pauli_I, pauli_mapping, pauli_X, pauli_Y, pauli_Z
from .skeletons import brickwall_skeleton, nn_edges
from .utils import endian_transform, pretty_print_sv, string_to_sv

__version__ = "0.1.0"


_MEASURE = ("partial_overlap", "rdm_from_sv", "measure_sv", "measure_and_collapse_sv", "pauli_expectation_sv",
            "marginal_probabilities")


def __getattr__(name):
    # the simulator (and with it the HIP library) loads on first use
    if name == "StatevectorSim":
        from .statevector_sim import StatevectorSim
        return StatevectorSim
    if name in _MEASURE:
        from . import measure
        return getattr(measure, name)
    raise AttributeError(name)
