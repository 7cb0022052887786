"""Dense 5-qubit blocks on the FP32 matrix cores vs the oracle (complex64,
tolerance ‖ψ − ψ_ref‖₂ ≤ 1e-5)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd.gates import haar_unitary
from qaravan_amd.workloads import haar_block_circuit

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [(0, 1, 2, 3, 4), (7, 8, 9, 10, 11), (3, 9, 1, 13, 6), (13, 12, 11, 10, 9),
                                  (2, 3, 4, 5, 6)])
def test_block5_direct_vs_oracle(gpu, bits):
    from qaravan_amd import _lib
    n = 14
    rng = np.random.default_rng(sum(bits))
    u = haar_unitary(32, rng)
    sv = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    sv /= np.linalg.norm(sv)
    psi = torch.tensor(sv, dtype=torch.complex64, device=gpu)
    ud = torch.tensor(u, dtype=torch.complex64, device=gpu)
    arr = (ctypes.c_int * 5)(*bits)
    lib = _lib.load()
    print(f"block5 bits={bits}", flush=True)
    _lib.check(lib.svh_apply_block5(_lib.ptr(psi), n, 0, arr, _lib.ptr(ud), _lib.stream_ptr(gpu)), "block5")
    # row bit j <-> address bit bits[j] <-> qubit n-1-bits[j]; qubit list is MSB first
    qubits = [n - 1 - bits[4 - a] for a in range(5)]
    ref = O.op_action(u, qubits, sv)
    err = np.linalg.norm(psi.cpu().numpy() - ref)
    print(f"  err={err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("n,layers", [(12, 3), (15, 6)])
def test_haar_block_circuit_vs_oracle(gpu, n, layers):
    from qaravan_amd import StatevectorSim
    circ = haar_block_circuit(n, layers)
    ref = O.run(circ.gate_list, n)
    sim = StatevectorSim(circ, device=gpu, dtype=torch.complex64)
    sim.run(progress_bar=False)
    err = np.linalg.norm(sim.get_statevector().cpu().numpy() - ref)
    sim128 = StatevectorSim(circ, device=gpu, dtype=torch.complex128)
    sim128.run(progress_bar=False)
    err128 = np.linalg.norm(sim128.get_statevector().cpu().numpy() - ref)
    print(f"haar n={n} layers={layers}: c64 err={err:.2e} c128 err={err128:.2e}", flush=True)
    assert err < 1e-5
    assert err128 < 1e-12


def test_block5_rejects_bad_arguments(gpu):
    from qaravan_amd import _lib
    lib = _lib.load()
    psi = torch.zeros(2 ** 12, dtype=torch.complex64, device=gpu)
    ud = torch.zeros(32 * 32, dtype=torch.complex64, device=gpu)
    assert lib.svh_apply_block5(_lib.ptr(psi), 12, 0, (ctypes.c_int * 5)(0, 1, 2, 3, 3), _lib.ptr(ud), None) == -4
    assert lib.svh_apply_block5(_lib.ptr(psi), 12, 1, (ctypes.c_int * 5)(0, 1, 2, 3, 4), _lib.ptr(ud), None) == -3
    assert lib.svh_apply_block5(_lib.ptr(psi), 11, 0, (ctypes.c_int * 5)(0, 1, 2, 3, 4), _lib.ptr(ud), None) == -2
