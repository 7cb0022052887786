"""ctypes binding of libqaravan_hip.so (include/qaravan_hip.h).

The library is the product path: there is no fallback.  If it is missing, or no
GPU is visible, calls raise ``RuntimeError`` instead of computing anything on
the CPU.  torch is imported first so the library binds to the same HIP runtime
instance torch uses (both carry the soname libamdhip64.so.7).
"""
import ctypes
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime the library links against)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqaravan_hip.so")

SVH_C64, SVH_C128 = 0, 1

_c_void_p = ctypes.c_void_p
_c_int = ctypes.c_int
_c_size_t = ctypes.c_size_t
_c_u64 = ctypes.c_uint64
_c_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); the exported C ABI, in header order
SIGNATURES = {
    "svh_last_error": (ctypes.c_char_p, []),
    "svh_abi_version": (_c_int, []),
    "svh_init_basis": (_c_int, [_c_void_p, _c_int, _c_int, _c_u64, _c_void_p]),
    "svh_apply_dense": (_c_int, [_c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_int, _c_void_p,
                                 _c_void_p]),
    "svh_apply_block5": (_c_int, [_c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_void_p, _c_void_p]),
    "svh_apply_block5x2": (_c_int, [_c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_void_p,
                                    ctypes.POINTER(_c_int), _c_void_p, _c_void_p]),
    "svh_permute_bits": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_void_p]),
    "svh_program_scratch_bytes": (_c_int, [_c_void_p, _c_int, ctypes.POINTER(_c_size_t)]),
    "svh_run_program": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                 _c_size_t, _c_void_p]),
    "svh_run_program_static": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "svh_sp_load": (_c_int, [ctypes.POINTER(ctypes.c_char_p), _c_int, _c_int, ctypes.POINTER(_c_void_p)]),
    "svh_sp_free": (_c_int, [_c_void_p]),
    "svh_run_program_sp": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p,
                                    _c_void_p, _c_size_t, _c_void_p]),
    "svh_run_adjoint_sp": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p,
                                    _c_void_p, _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "svh_run_adjoint": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p,
                                 _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "svh_norm2": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "svh_scale_by_inv_norm": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p]),
    "svh_pauli_expect": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p,
                                  _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "svh_pauli_apply": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_void_p,
                                 _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "svh_pauli_apply_table": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_void_p]),
    "svh_pauli_apply_inner": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_void_p, _c_void_p,
                                       _c_void_p, _c_size_t, _c_void_p]),
    "svh_inner": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_size_t,
                           _c_void_p]),
    "svh_batch_tile": (_c_int, [_c_void_p, _c_int, ctypes.c_int64, _c_int, _c_int, _c_int, _c_int, _c_void_p]),
    "svh_batch_reduce": (_c_int, [_c_void_p, _c_int, ctypes.c_int64, _c_void_p]),
    "svh_batch_hamiltonian": (_c_int, [_c_void_p, _c_int, ctypes.c_int64, _c_int, _c_void_p]),
    "svh_batch_init": (_c_int, [_c_void_p, ctypes.c_int64, _c_void_p, _c_int, _c_int, _c_void_p]),
    "svh_batch_trig": (_c_int, [_c_void_p, ctypes.c_int64, _c_void_p, _c_int, _c_void_p]),
    "svh_batch_persistent": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_int,
                                      _c_void_p, _c_int, _c_size_t, _c_int, _c_int, _c_int, ctypes.POINTER(_c_int),
                                      _c_void_p, _c_void_p]),
    "svh_timestamp": (_c_int, [_c_void_p, _c_void_p]),
    "svh_spb_load": (_c_int, [ctypes.POINTER(ctypes.c_char_p), _c_int, ctypes.POINTER(ctypes.c_char_p), _c_int,
                              ctypes.POINTER(_c_void_p)]),
    "svh_spb_launch": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_int, ctypes.c_int64, _c_int, _c_size_t,
                                _c_void_p]),
    "svh_measure_scratch_bytes": (_c_size_t, [_c_int, _c_int, _c_int]),
    "svh_marginal_probs": (_c_int, [_c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_int, _c_void_p,
                                    _c_void_p, _c_size_t, _c_void_p]),
    "svh_sample_cdf": (_c_int, [_c_void_p, _c_int, _c_void_p, ctypes.c_int64, _c_void_p, _c_void_p, _c_void_p,
                                _c_size_t, _c_void_p]),
    "svh_collapse_gather": (_c_int, [_c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_int, _c_u64,
                                     _c_void_p, _c_void_p]),
    "svh_partial_overlap": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, ctypes.POINTER(_c_int), _c_int,
                                     _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
}

SVH_MEAS_PROBS, SVH_MEAS_SAMPLE, SVH_MEAS_OVERLAP = 1, 2, 3

_lib = None
_lock = threading.Lock()


class HipLibraryError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the library (once) and declare every entry point's signature."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise HipLibraryError(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the HIP path has no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.svh_abi_version() != 1:
            raise HipLibraryError("libqaravan_hip.so ABI version mismatch; rebuild it")
        _lib = lib
        return lib


def check(status, what=""):
    """Map an ABI status to an exception (<0 argument error, >0 HIP error)."""
    if status == 0:
        return
    msg = load().svh_last_error().decode(errors="replace")
    if status < 0:
        raise ValueError(f"{what}: {msg} (status {status})")
    raise HipLibraryError(f"{what}: HIP error {status}: {msg}")


def require_gpu():
    if not torch.cuda.is_available():
        raise HipLibraryError("no GPU visible: the qaravan_amd statevector path runs only on the GPU")


def dtype_code(dtype):
    if dtype == torch.complex64:
        return SVH_C64
    if dtype == torch.complex128:
        return SVH_C128
    raise ValueError(f"state dtype must be complex64 or complex128, got {dtype}")


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())
