"""Writes the fixtures in this folder from the oracle (complex128).

brickwork_12q_20l_state.npz: final state of the 12-qubit, 20-layer seeded
brickwork circuit (830 gates) from |0…0⟩.  Regenerate with
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import sv_oracle as O  # noqa: E402
from qaravan_amd.workloads import brickwork_circuit  # noqa: E402


def main():
    circ = brickwork_circuit(12, 20)
    sv = O.run(circ.gate_list, 12)
    np.savez_compressed(os.path.join(HERE, "brickwork_12q_20l_state.npz"), state=sv)
    print("wrote brickwork_12q_20l_state.npz", np.linalg.norm(sv))


if __name__ == "__main__":
    main()
