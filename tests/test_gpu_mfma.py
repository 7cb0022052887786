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


@pytest.mark.parametrize("n,ba,bb", [
    (14, (0, 1, 2, 3, 4), (5, 6, 7, 8, 9)),          # one contiguous 4096-amplitude tile
    (20, (10, 11, 12, 13, 14), (15, 16, 17, 18, 19)),  # 32-byte runs, sibling tiles grouped per XCD
    (20, (15, 16, 17, 18, 19), (2, 3, 4, 5, 6)),      # lowest other bits 0, 1; no sibling grouping
    (20, (3, 4, 5, 6, 7), (15, 16, 17, 18, 19)),      # three spare bits (0, 1, 2): 64-byte runs
    (16, (3, 4, 5, 6, 7), (11, 12, 13, 14, 15)),      # three spare bits, whole-line runs
    (13, (8, 9, 10, 11, 12), (3, 4, 5, 6, 7)),        # three spare bits, a single 8192-amplitude tile
    (17, (3, 9, 1, 13, 6), (16, 0, 11, 2, 8)),        # scrambled bits
    (12, (11, 10, 9, 8, 7), (6, 5, 4, 3, 2)),         # single tile
    (24, (10, 11, 12, 13, 14), (15, 16, 17, 18, 19)),  # 8 tiles per workgroup: the prefetch pipeline
    (23, (3, 4, 5, 6, 7), (18, 19, 20, 21, 22)),      # 4 tiles per workgroup, three spare bits
])
def test_block5x2_direct_vs_oracle(gpu, n, ba, bb):
    from qaravan_amd import _lib
    rng = np.random.default_rng(sum(ba) + 7 * sum(bb))
    ua, ub = haar_unitary(32, rng), haar_unitary(32, rng)
    sv = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    sv /= np.linalg.norm(sv)
    psi = torch.tensor(sv, dtype=torch.complex64, device=gpu)
    uad = torch.tensor(ua, dtype=torch.complex64, device=gpu)
    ubd = torch.tensor(ub, dtype=torch.complex64, device=gpu)
    lib = _lib.load()
    print(f"block5x2 n={n} a={ba} b={bb}", flush=True)
    _lib.check(lib.svh_apply_block5x2(_lib.ptr(psi), n, 0, (ctypes.c_int * 5)(*ba), _lib.ptr(uad),
                                      (ctypes.c_int * 5)(*bb), _lib.ptr(ubd), _lib.stream_ptr(gpu)), "block5x2")
    ref = O.op_action(ua, [n - 1 - ba[4 - a] for a in range(5)], sv)
    ref = O.op_action(ub, [n - 1 - bb[4 - a] for a in range(5)], ref)
    err = np.linalg.norm(psi.cpu().numpy() - ref)
    print(f"  err={err:.2e}", flush=True)
    assert err < 1e-5


def test_block5x2_rejects_bad_arguments(gpu):
    from qaravan_amd import _lib
    lib = _lib.load()
    psi = torch.zeros(2 ** 12, dtype=torch.complex64, device=gpu)
    ud = torch.zeros(32 * 32, dtype=torch.complex64, device=gpu)
    a = (ctypes.c_int * 5)(0, 1, 2, 3, 4)
    assert lib.svh_apply_block5x2(_lib.ptr(psi), 12, 0, a, _lib.ptr(ud), (ctypes.c_int * 5)(4, 5, 6, 7, 8),
                                  _lib.ptr(ud), None) == -4
    assert lib.svh_apply_block5x2(_lib.ptr(psi), 12, 1, a, _lib.ptr(ud), (ctypes.c_int * 5)(5, 6, 7, 8, 9),
                                  _lib.ptr(ud), None) == -3
    assert lib.svh_apply_block5x2(_lib.ptr(psi), 11, 0, a, _lib.ptr(ud), (ctypes.c_int * 5)(5, 6, 7, 8, 9),
                                  _lib.ptr(ud), None) == -2


@pytest.mark.parametrize("n,layers", [(12, 3), (15, 6), (20, 5), (17, 12)])
def test_haar_block_circuit_vs_oracle(gpu, n, layers):
    from qaravan_amd import StatevectorSim
    circ = haar_block_circuit(n, layers)
    ref = O.run(circ.gate_list, n)
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=torch.complex64)
    sim.run(progress_bar=False)
    err = np.linalg.norm(sim.get_statevector().cpu().numpy() - ref)
    sim128 = StatevectorSim(circ, backend="hip", device=gpu, dtype=torch.complex128)
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
