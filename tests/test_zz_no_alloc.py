"""The allocation-free program runner (svh_run_program_static) against the
caller-scratch runner (svh_run_program): same program, same input, so the
results must be identical bit for bit.  Two calls in a row, the second on a
smaller register, and a program whose scratch exceeds the static buffer,
which must be refused."""
import ctypes

import numpy as np
import pytest
import torch

from qaravan_amd import _lib
from qaravan_amd import program as P
from qaravan_amd.statevector_sim import DeviceScratch, run_program
from qaravan_amd.workloads import hea_ops, hea_theta
from test_gpu_parity import random_circuit

pytestmark = pytest.mark.gpu


def _static_run(psi, prog, theta=None):
    lib = _lib.load()
    dev = prog.device_copy(psi.device)
    th = None if theta is None else _lib.ptr(theta)
    _lib.check(lib.svh_run_program_static(_lib.ptr(psi), prog.n, _lib.dtype_code(psi.dtype), prog.blob.ctypes.data,
                                          _lib.ptr(dev), th, _lib.stream_ptr(psi.device)), "svh_run_program_static")


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_static_scratch_matches_caller_scratch(gpu, dtype):
    rng = np.random.default_rng(77)
    for n, cfg in ((16, P.TileConfig(tb=7, low_bits=4, max_cost=96, rb=3, generic=True, io_lane_bits=4)),
                   (12, P.TileConfig(tb=6, low_bits=4, max_cost=48))):
        circ = random_circuit(n, 150, rng, k3=True).build()
        prog = P.compile_ops(P.lower_circuit(circ, n), n, cfg)
        _compare(gpu, prog, dtype, rng)
    # device angles: cos/sin tables live in the scratch too
    n, layers = 14, 4
    cfg = P.TileConfig(tb=8, low_bits=4, max_cost=1e9, rb=3, generic=True, io_lane_bits=4)
    prog = P.compile_ops(hea_ops(n, layers), n, cfg, n_params=2 * n * layers)
    theta = torch.tensor(hea_theta(n, layers), dtype=torch.float64, device=gpu)
    _compare(gpu, prog, dtype, rng, theta)


def _compare(gpu, prog, dtype, rng, theta=None):
    n = prog.n
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    a = torch.tensor(v, dtype=dtype, device=gpu)
    b = a.clone()
    print(f"static scratch n={n} {dtype} {prog.describe()}", flush=True)
    run_program(a, prog, DeviceScratch(gpu), theta)
    _static_run(b, prog, theta)
    torch.cuda.synchronize(gpu)
    assert torch.equal(a, b)
    assert not torch.equal(a, torch.tensor(v, dtype=dtype, device=gpu))


def test_static_scratch_refuses_large_programs(gpu):
    # a remapped program ends with a permutation through a state-sized
    # buffer: at 22 qubits that is 32 MiB, more than the static scratch
    n = 22
    cfg = P.TileConfig(tb=8, low_bits=4, max_cost=1e9, rb=3, generic=True, io_lane_bits=4, remap=True)
    rng = np.random.default_rng(78)
    circ = random_circuit(n, 60, rng, k3=False).build()
    prog = P.compile_ops(P.lower_circuit(circ, n), n, cfg)
    assert prog.final_pos is not None
    psi = torch.zeros(2 ** n, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):
        _static_run(psi, prog)
