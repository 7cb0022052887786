"""Measurement, sampling, reduced density matrices and partial overlaps on the GPU.

Mirrors the module functions of src/qaravan/tensorQ/statevector_sim.py:

* ``partial_overlap(sv1, sv2, local_dim=2, skip=None)`` (:155-180):
  M[i, j] = Σ_rest sv1[i, rest]·conj(sv2[j, rest]) over the kept (``skip``)
  sites in ascending order, big-endian; a (2^k, 2^k) complex128 tensor.
* ``rdm_from_sv(sv, sites, local_dim=2)`` (:182-183).
* ``measure_sv(sv, meas_sites, shots=1)`` (:185-193): a bitstring (or a list
  of ``shots`` of them) over the SORTED sites, drawn like
  ``np.random.choice(outcomes, p=probs)``: the uniforms come from numpy's
  global RNG on the host (``np.random.random_sample``), so a seeded
  ``np.random.seed(s)`` run gives the reference's outcomes.
* ``measure_and_collapse_sv(sv, meas_sites, local_dim=2)`` (:195-215): the
  outcome and the normalised state of the remaining sites.  As in the
  reference, outcome character j is projected onto site ``meas_sites[j]``
  (this matches the drawn outcome when ``meas_sites`` is sorted).
* ``pauli_expectation_sv(sv, pauli_string)`` (:217-234).

The probability vector is the diagonal of the reduced density matrix, computed
directly as Σ_rest |ψ|² without forming the 2^k x 2^k matrix, so measuring
many sites of a large register is one pass over the state.  Every sum is
accumulated in FP64 in a fixed order.  States are device tensors (complex64 or
complex128, flat or rank-n); host arrays are copied to the current GPU.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .statevector_sim import DeviceScratch, pauli_sum_expectation

_scratch = {}


def _scratch_for(device):
    key = str(device)
    if key not in _scratch:
        _scratch[key] = DeviceScratch(device)
    return _scratch[key]


def _as_state(sv, device=None):
    """Flat contiguous device tensor and its qubit count."""
    if not torch.is_tensor(sv):
        sv = torch.as_tensor(np.asarray(sv))
    if not sv.is_complex():
        sv = sv.to(torch.complex128)
    if sv.device.type != "cuda":
        _lib.require_gpu()
        sv = sv.to(device or "cuda")
    flat = sv.reshape(-1)
    if not flat.is_contiguous():
        flat = flat.contiguous()
    N = flat.numel()
    n = N.bit_length() - 1
    if N < 2 or (1 << n) != N:
        raise ValueError(f"state has {N} amplitudes, not a power of two >= 2")
    _lib.dtype_code(flat.dtype)
    return flat, n


def _site_bits(n, sites, sort):
    sites = [int(q) for q in sites]
    if len(set(sites)) != len(sites):
        raise ValueError(f"repeated site in {sites}")
    for q in sites:
        if not 0 <= q < n:
            raise ValueError(f"site {q} outside a {n}-qubit register")
    if sort:
        sites = sorted(sites)
    arr = (ctypes.c_int * max(1, len(sites)))(*[n - 1 - q for q in sites])
    return sites, arr


def marginal_probabilities(sv, sites):
    """p[m] = Σ_rest |ψ(m, rest)|² over the sorted ``sites`` (float64, 2^k)."""
    flat, n = _as_state(sv)
    sites, bits = _site_bits(n, sites, sort=True)
    return _marginal(flat, n, bits, len(sites))


def _marginal(flat, n, bits, k):
    lib = _lib.load()
    out = torch.empty(1 << k, dtype=torch.float64, device=flat.device)
    need = lib.svh_measure_scratch_bytes(_lib.SVH_MEAS_PROBS, n, k)
    buf = _scratch_for(flat.device).get(need)
    _lib.check(lib.svh_marginal_probs(_lib.ptr(flat), n, _lib.dtype_code(flat.dtype), bits, k, _lib.ptr(out),
                                      _lib.ptr(buf), buf.numel(), _lib.stream_ptr(flat.device)),
               "svh_marginal_probs")
    return out


def _sum_tolerance(dtype):
    # numpy's np.random.choice rejects p whose sum is off by more than
    # sqrt(float64 eps); a complex64 state carries FP32 rounding in its norm.
    return float(np.sqrt(np.finfo(np.float64).eps)) if dtype == torch.complex128 else 1e-4


def sample_indices(probs, shots, state_dtype=torch.complex128):
    """Indices drawn like np.random.choice(len(probs), p=probs, size=shots),
    with the uniforms taken from numpy's global RNG.  ``shots=None`` draws
    one scalar (size=None)."""
    lib = _lib.load()
    M = probs.numel()
    k = M.bit_length() - 1
    if (1 << k) != M:
        raise ValueError("number of outcomes must be a power of two")
    dev = probs.device
    need = lib.svh_measure_scratch_bytes(_lib.SVH_MEAS_SAMPLE, max(k, 1), k)
    buf = _scratch_for(dev).get(need)
    total = torch.empty(1, dtype=torch.float64, device=dev)
    st = _lib.stream_ptr(dev)
    _lib.check(lib.svh_sample_cdf(_lib.ptr(probs), k, None, 0, None, _lib.ptr(total), _lib.ptr(buf),
                                  buf.numel(), st), "svh_sample_cdf")
    tot = float(total.item())
    if not np.isfinite(tot):
        raise ValueError("probabilities contain NaN")
    if abs(tot - 1.0) > _sum_tolerance(state_dtype):
        raise ValueError("probabilities do not sum to 1")
    count = 1 if shots is None else int(shots)
    u = np.random.random_sample() if shots is None else np.random.random_sample(count)
    u_dev = torch.as_tensor(np.atleast_1d(np.asarray(u, dtype=np.float64))).to(dev)
    idx = torch.empty(count, dtype=torch.int64, device=dev)
    _lib.check(lib.svh_sample_cdf(_lib.ptr(probs), k, _lib.ptr(u_dev), count, _lib.ptr(idx), _lib.ptr(total),
                                  _lib.ptr(buf), buf.numel(), st), "svh_sample_cdf")
    return idx.cpu().numpy()


def measure_sv(sv, meas_sites, shots=1):
    """Bitstring outcome(s) of measuring ``meas_sites`` (statevector_sim.py:185-193)."""
    flat, n = _as_state(sv)
    sites, bits = _site_bits(n, meas_sites, sort=True)
    k = len(sites)
    probs = _marginal(flat, n, bits, k)
    idx = sample_indices(probs, None if shots == 1 else shots, flat.dtype)
    outs = _bitstrings(idx, k)
    return outs[0] if shots == 1 else outs


def _bitstrings(idx, k):
    """np.binary_repr(i, width=k) for every index (same strings, faster)."""
    if k == 0:
        return [np.binary_repr(0, width=0)] * len(idx)
    return list(map(("{:0%db}" % k).format, np.asarray(idx, dtype=np.int64).tolist()))


def measure_and_collapse_sv(sv, meas_sites, local_dim=2):
    """(outcome, normalised post-measurement state of the other sites)
    (statevector_sim.py:195-215)."""
    if local_dim != 2:
        raise NotImplementedError("the HIP statevector path supports qubits (local_dim = 2) only")
    flat, n = _as_state(sv)
    outcome = measure_sv(flat, meas_sites)
    sites, bits = _site_bits(n, meas_sites, sort=False)
    k = len(sites)
    lib = _lib.load()
    code = _lib.dtype_code(flat.dtype)
    st = _lib.stream_ptr(flat.device)
    out = torch.empty(1 << (n - k), dtype=flat.dtype, device=flat.device)
    _lib.check(lib.svh_collapse_gather(_lib.ptr(flat), n, code, bits, k, int(outcome, 2) if k else 0,
                                       _lib.ptr(out), st), "svh_collapse_gather")
    norm2 = torch.empty(1, dtype=torch.float64, device=flat.device)
    buf = _scratch_for(flat.device).get(8 * 1024)
    _lib.check(lib.svh_norm2(_lib.ptr(out), n - k, code, _lib.ptr(norm2), _lib.ptr(buf), buf.numel(), st),
               "svh_norm2")
    _lib.check(lib.svh_scale_by_inv_norm(_lib.ptr(out), n - k, code, _lib.ptr(norm2), st),
               "svh_scale_by_inv_norm")
    return outcome, out


def partial_overlap(sv1, sv2, local_dim=2, skip=None):
    """(2^k, 2^k) complex128 M[i, j] = Σ_rest sv1[i, rest]·conj(sv2[j, rest])
    over the sorted ``skip`` sites (statevector_sim.py:155-180); at most 6
    kept sites."""
    if local_dim != 2:
        raise NotImplementedError("the HIP statevector path supports qubits (local_dim = 2) only")
    a, n = _as_state(sv1)
    b, n2 = _as_state(sv2, a.device)
    if n2 != n:
        raise ValueError("states have different sizes")
    if b.dtype != a.dtype:
        b = b.to(a.dtype)
    if b.device != a.device:
        b = b.to(a.device)
    sites, bits = _site_bits(n, skip or [], sort=True)
    k = len(sites)
    lib = _lib.load()
    need = lib.svh_measure_scratch_bytes(_lib.SVH_MEAS_OVERLAP, n, k)
    if need == 0:
        raise ValueError("partial_overlap keeps at most 6 sites on the HIP path")
    buf = _scratch_for(a.device).get(need)
    out = torch.empty((1 << k, 1 << k), dtype=torch.complex128, device=a.device)
    _lib.check(lib.svh_partial_overlap(_lib.ptr(a), _lib.ptr(b), n, _lib.dtype_code(a.dtype), bits, k,
                                       _lib.ptr(out), _lib.ptr(buf), buf.numel(), _lib.stream_ptr(a.device)),
               "svh_partial_overlap")
    return out


def rdm_from_sv(sv, sites, local_dim=2):
    """Reduced density matrix of ``sites`` (statevector_sim.py:182-183)."""
    return partial_overlap(sv, sv, local_dim=local_dim, skip=sites)


def pauli_expectation_sv(sv, pauli_string):
    """Re⟨ψ|P|ψ⟩ for a string over 'ixyz' (statevector_sim.py:217-234)."""
    flat, n = _as_state(sv)
    if len(pauli_string) != n:
        raise AssertionError("Pauli string length must match number of qubits in statevector")
    return float(pauli_sum_expectation(flat, n, [(1.0, pauli_string)], _scratch_for(flat.device)).item())
