"""Qubit remapping on the GPU: the out-of-place bit permutation kernel
against a numpy transpose, and remapped programs (each pass stores its tile
with permuted bits, the run permutes back) against the oracle."""
import ctypes

import numpy as np
import pytest
import torch

import prog_interp
from oracle import sv_oracle as O
from qaravan_amd import program as P
from qaravan_amd.workloads import brickwork_circuit

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed", [(8, 0), (11, 1), (14, 2), (18, 3)])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_permute_bits_vs_numpy(gpu, n, seed, dtype):
    from qaravan_amd import _lib
    rng = np.random.default_rng(seed)
    perm = [int(x) for x in rng.permutation(n)]
    sv = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    src = torch.tensor(sv, dtype=dtype, device=gpu)
    dst = torch.empty_like(src)
    print(f"permute n={n} perm={perm} {dtype}", flush=True)
    lib = _lib.load()
    _lib.check(lib.svh_permute_bits(_lib.ptr(src), _lib.ptr(dst), n, _lib.dtype_code(dtype),
                                    (ctypes.c_int * n)(*perm), _lib.stream_ptr(gpu)), "svh_permute_bits")
    ref = prog_interp.permute_bits(sv.astype(dst.cpu().numpy().dtype), n, {b: perm[b] for b in range(n)})
    assert np.array_equal(dst.cpu().numpy(), ref)


def test_permute_bits_rejects_bad_arguments(gpu):
    from qaravan_amd import _lib
    lib = _lib.load()
    a = torch.zeros(2 ** 8, dtype=torch.complex64, device=gpu)
    b = torch.zeros_like(a)
    ok = (ctypes.c_int * 8)(*range(8))
    assert lib.svh_permute_bits(_lib.ptr(a), _lib.ptr(a), 8, 0, ok, None) == -1
    assert lib.svh_permute_bits(_lib.ptr(a), _lib.ptr(b), 8, 0, (ctypes.c_int * 8)(0, 1, 2, 3, 4, 5, 6, 6), None) == -4
    assert lib.svh_permute_bits(_lib.ptr(a), _lib.ptr(b), 7, 0, ok, None) == -2


@pytest.mark.parametrize("n,layers,tb,cap,split", [(14, 8, 6, 1e9, 1), (16, 10, 6, 96, 0), (18, 12, 8, 160, 1),
                                                   (20, 14, 8, 1e9, 2)])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_remapped_brickwork_vs_oracle(gpu, n, layers, tb, cap, split, dtype):
    from qaravan_amd import StatevectorSim
    if dtype == torch.complex128 and tb + 5 - split > 13:
        pytest.skip("c128 tile too large for LDS")
    circ = brickwork_circuit(n, layers)
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=cap, specialize=True, io_lane_bits=4, rb=5, lds_split=split,
                       remap=True)
    print(f"remap n={n} layers={layers} cfg={cfg} {dtype}", flush=True)
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=dtype, tile=cfg)
    sim.run(progress_bar=False)
    assert sim._prog.final_pos is not None
    sv = sim.get_statevector().cpu().numpy()
    ref = O.run(circ.gate_list, n)
    err = np.linalg.norm(sv - ref)
    print(f"  err={err:.2e}", flush=True)
    assert err < (1e-12 if dtype == torch.complex128 else 1e-5)


def test_remapped_brickwork_24q_c64_vs_c128_and_inverse(gpu):
    from qaravan_amd import Circuit, StatevectorSim
    circ = brickwork_circuit(24, 12)
    cfg = P.TileConfig(tb=8, low_bits=4, max_cost=160, specialize=True, io_lane_bits=4, rb=5, lds_split=1, remap=True)
    ref = StatevectorSim(circ, backend="hip", device=gpu, dtype=torch.complex128, tile=cfg)
    ref.run(progress_bar=False)
    a = ref.get_statevector().cpu().numpy()
    sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=torch.complex64, tile=cfg)
    sim.run(progress_bar=False)
    b = sim.get_statevector().cpu().numpy()
    assert abs(np.linalg.norm(a) - 1) < 1e-12
    assert np.linalg.norm(a - b) < 1e-5
    inv = Circuit(circ.gate_list + [g.dag() for g in circ.gate_list[::-1]], n=24)
    z = StatevectorSim(inv, backend="hip", device=gpu, dtype=torch.complex128, tile=cfg)
    z.run(progress_bar=False)
    assert abs(z.get_statevector()[0].item() - 1) < 1e-11


@pytest.mark.parametrize("n,tb", [(14, 6), (16, 7)])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_trailing_cnots_in_final_permutation_vs_oracle(gpu, n, tb, dtype):
    """Remapped programs whose trailing CNOT pass is folded into the final
    permutation (the permutation kernel applies the CNOTs while it gathers)
    give the oracle's state."""
    from qaravan_amd import StatevectorSim
    from test_program import _trailing_cx_circuits
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=1e9, specialize=True, io_lane_bits=4, rb=5, lds_split=1,
                       remap=True)
    folded = 0
    for circ in _trailing_cx_circuits(n, range(6)):
        sim = StatevectorSim(circ, backend="hip", device=gpu, dtype=dtype, tile=cfg)
        sim.run(progress_bar=False)
        folded += bool(sim._prog.final_cx)
        err = np.linalg.norm(sim.get_statevector().cpu().numpy() - O.run(circ.gate_list, n))
        print(f"  trailing cnots n={n} final_cx={sim._prog.final_cx} err={err:.2e}", flush=True)
        assert err < (1e-12 if dtype == torch.complex128 else 1e-5)
    assert folded > 0
