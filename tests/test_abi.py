"""The C-ABI library loads, exports what the header declares, and rejects bad
arguments before touching the GPU (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from qaravan_amd import _lib
from qaravan_amd import program as P
from qaravan_amd.workloads import brickwork_circuit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(svh_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) <= set(_lib.SIGNATURES), "ctypes signatures missing for: %s" % (set(syms) - set(_lib.SIGNATURES))


def test_abi_version():
    assert _lib.load().svh_abi_version() == 1


def test_argument_errors_without_gpu():
    lib = _lib.load()
    assert lib.svh_init_basis(None, 3, 0, 0, None) == -1
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: rejected before any launch
    assert lib.svh_init_basis(fake, 0, 0, 0, None) == -2
    assert lib.svh_init_basis(fake, 40, 0, 0, None) == -2
    assert lib.svh_init_basis(fake, 3, 7, 0, None) == -3
    assert lib.svh_init_basis(fake, 3, 0, 8, None) == -4
    bits = (ctypes.c_int * 2)(0, 0)
    assert lib.svh_apply_dense(fake, 3, 0, bits, 2, fake, None) == -4   # repeated bit
    bits = (ctypes.c_int * 1)(5)
    assert lib.svh_apply_dense(fake, 3, 0, bits, 1, fake, None) == -4   # bit out of range
    assert b"out of range" in lib.svh_last_error()


def test_measurement_argument_errors_without_gpu():
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: rejected before any launch
    rep = (ctypes.c_int * 2)(3, 3)
    assert lib.svh_marginal_probs(fake, 6, 0, rep, 2, fake, fake, 1 << 20, None) == -2
    assert b"repeated" in lib.svh_last_error()
    far = (ctypes.c_int * 1)(6)
    assert lib.svh_marginal_probs(fake, 6, 0, far, 1, fake, fake, 1 << 20, None) == -2
    seven = (ctypes.c_int * 7)(*range(7))
    assert lib.svh_partial_overlap(fake, fake, 10, 0, seven, 7, fake, fake, 1 << 30, None) == -2
    two = (ctypes.c_int * 2)(0, 5)
    assert lib.svh_collapse_gather(fake, 6, 0, two, 2, 4, fake, None) == -4   # outcome beyond 2 bits
    assert lib.svh_partial_overlap(fake, fake, 10, 0, two, 2, fake, fake, 16, None) == -5  # scratch
    assert lib.svh_sample_cdf(fake, 3, None, 5, None, fake, fake, 1 << 20, None) == -1
    assert lib.svh_measure_scratch_bytes(3, 10, 7) == 0
    assert lib.svh_measure_scratch_bytes(2, 10, 10) >= 8 * 1024
    assert lib.svh_measure_scratch_bytes(9, 10, 2) == 0


def test_program_validation_on_host():
    lib = _lib.load()
    prog = P.compile_ops(P.lower_circuit(brickwork_circuit(12, 2).build(), 12), 12, P.TileConfig(tb=6))
    need = ctypes.c_size_t()
    assert lib.svh_program_scratch_bytes(prog.blob.ctypes.data, 0, ctypes.byref(need)) == 0
    assert need.value >= 8 * 1024
    bad = prog.blob.copy()
    bad[0] ^= 0xFF
    assert lib.svh_program_scratch_bytes(bad.ctypes.data, 0, ctypes.byref(need)) == -8
    # corrupt a register bit of the first op
    bad = prog.blob.copy()
    h = np.frombuffer(bad[:96].tobytes(), dtype=P.HDR_DT)[0]
    off = int(h['off_ops'])
    ops = bad[off:off + P.OP_DT.itemsize].view(P.OP_DT)
    ops[0]['a'] = 9
    assert lib.svh_program_scratch_bytes(bad.ctypes.data, 0, ctypes.byref(need)) == -8
    assert b"register bit" in lib.svh_last_error()


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from qaravan_amd import StatevectorSim
    from qaravan_amd.workloads import brickwork_circuit as bw
    with pytest.raises(RuntimeError):
        StatevectorSim(bw(4, 1))                      # numpy results, GPU compute: no GPU here
    with pytest.raises(RuntimeError):
        StatevectorSim(bw(4, 1), backend="torch")
    with pytest.raises(ValueError):
        StatevectorSim(bw(4, 1), backend="jax")
