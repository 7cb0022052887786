"""Autotuner for the fused tile passes.

Picks (tb, low_bits, max_cost) — workgroup size (tile = tb + 4 bits), the
number of low address bits every tile keeps (contiguous run length), and the
per-pass ALU cap — by timing the compiled program of a representative circuit
on the GPU.  Results are cached per (n, dtype) in tune_cache.json next to this
file; ``best_tile_config`` falls back to the built-in defaults when no entry
exists.
"""
import itertools
import json
import os
import time
from dataclasses import replace

from .program import TileConfig

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tune_cache.json")

DEFAULTS = {"complex64": TileConfig(tb=8, low_bits=6, max_cost=96.0),
            "complex128": TileConfig(tb=8, low_bits=6, max_cost=96.0)}


def _key(n, dtype):
    return f"{str(dtype).replace('torch.', '')}:{n}"


def _load():
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def best_tile_config(n, dtype):
    cache = _load()
    name = str(dtype).replace('torch.', '')
    e = cache.get(_key(n, dtype))
    if e is None:
        # nearest tuned size for this dtype
        cands = [(abs(int(k.split(':')[1]) - n), v) for k, v in cache.items() if k.startswith(name + ':')]
        if cands:
            e = min(cands, key=lambda t: t[0])[1]
    base = DEFAULTS.get(name, TileConfig())
    if e is None:
        cfg = base
    else:
        cfg = TileConfig(tb=e["tb"], low_bits=e["low_bits"], max_cost=e["max_cost"],
                         prefetch=e.get("prefetch", False), specialize=e.get("specialize", False),
                         fuse2q=e.get("fuse2q", True), io_lane_bits=e.get("io_lane_bits", 6), rb=e.get("rb", 4),
                         lds_split=e.get("lds_split", 0), remap=e.get("remap", False),
                         remap_free=e.get("remap_free", -1))
    return replace(cfg, tb=min(cfg.tb, n - cfg.rb))


# below this size a run is launch- or latency-bound and the per-pass compiled
# kernels (hipcc at first use of each new circuit structure) do not pay
SIM_TUNED_MIN_QUBITS = 20


def sim_tile_config(n, dtype):
    """The tile configuration ``StatevectorSim`` uses when none is given: the
    tuned one (nearest tuned size of this dtype) from SIM_TUNED_MIN_QUBITS
    qubits up when the cache holds an entry for the dtype, else the
    op-interpreting defaults, which need no compiler at run time."""
    name = str(dtype).replace('torch.', '')
    if n >= SIM_TUNED_MIN_QUBITS and any(k.startswith(name + ':') for k in _load()):
        return best_tile_config(n, dtype)
    return DEFAULTS.get(name, TileConfig())


def tune(circ, dtype, device, space=None, repeats=3, write=True, log=print, out_path=None):
    """Time every configuration in ``space`` on ``circ`` and cache the fastest."""
    import torch

    from . import _lib
    from .program import compile_ops, lower_circuit
    from .statevector_sim import DeviceScratch, run_program
    n = circ.num_sites
    space = space or dict(tb=[6, 7, 8, 9], low_bits=[3, 4, 5, 6], max_cost=[48.0, 96.0, 160.0, 1e9],
                          prefetch=[False, True])
    lib = _lib.load()
    ops = lower_circuit(circ.build(), n)
    psi = torch.empty(2 ** n, dtype=dtype, device=device)
    scratch = DeviceScratch(device)
    results = []
    for tb, lb, mc, pf, spz, f2, io, rb, hl, rm, rf in itertools.product(
            space["tb"], space["low_bits"], space["max_cost"], space.get("prefetch", [False]),
            space.get("specialize", [False]), space.get("fuse2q", [True]), space.get("io_lane_bits", [6]),
            space.get("rb", [4]), space.get("lds_split", [0]), space.get("remap", [False]),
            space.get("remap_free", [-1])):
        if tb + rb > n or (dtype == torch.complex128 and tb > (8 if pf else 9)) or (pf and tb > 9):
            continue
        if (pf and spz) or (rb == 5 and (pf or not spz)) or (hl and (pf or not spz or hl >= rb)) or \
                (rm and (pf or not spz)) or (not rm and rf != space.get("remap_free", [-1])[0]):
            continue
        cfg = TileConfig(tb=tb, low_bits=lb, max_cost=mc, prefetch=pf, specialize=spz, fuse2q=f2, io_lane_bits=io,
                         rb=rb, lds_split=hl, remap=rm, remap_free=rf)
        prog = compile_ops(ops, n, cfg)
        times = []
        for r in range(repeats + 1):
            _lib.check(lib.svh_init_basis(_lib.ptr(psi), n, _lib.dtype_code(dtype), 0, _lib.stream_ptr(device)),
                       "init")
            torch.cuda.synchronize(device)
            t0 = time.perf_counter()
            run_program(psi, prog, scratch)
            torch.cuda.synchronize(device)
            if r:
                times.append(time.perf_counter() - t0)
        t = min(times)
        results.append((t, cfg, prog.n_tile_passes()))
        log(f"[tune] n={n} {dtype} tb={tb} rb={rb} low={lb} cap={mc:g} prefetch={pf} specialize={spz} fuse2q={f2} "
            f"io={io} lds_split={hl} remap={rm}/{rf}: "
            f"{t * 1e3:.3f} ms, {prog.n_tile_passes()} passes")
    results.sort(key=lambda x: x[0])
    t, cfg, npass = results[0]
    if write:
        cache = _load()
        cache[_key(n, dtype)] = dict(tb=cfg.tb, low_bits=cfg.low_bits, max_cost=cfg.max_cost,
                                     prefetch=cfg.prefetch, specialize=cfg.specialize, fuse2q=cfg.fuse2q,
                                     io_lane_bits=cfg.io_lane_bits, rb=cfg.rb, lds_split=cfg.lds_split,
                                     remap=cfg.remap, remap_free=cfg.remap_free,
                                     ms=t * 1e3,
                                     passes=npass)
        with open(out_path or CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return cfg, results


ADJOINT_DEFAULT = {"complex64": TileConfig(tb=6, low_bits=4, max_cost=1e9, specialize=True),
                   "complex128": TileConfig(tb=6, low_bits=4, max_cost=1e9, specialize=True)}


def best_adjoint_config(n, dtype):
    """Tile configuration for forward+adjoint of parametrised circuits (cached
    under 'adjoint:<dtype>:<n>' when tuned, else the default)."""
    name = str(dtype).replace('torch.', '')
    e = _load().get(f"adjoint:{name}:{n}")
    if e is None:
        cfg = ADJOINT_DEFAULT[name]
    else:
        cfg = TileConfig(tb=e["tb"], low_bits=e["low_bits"], max_cost=e["max_cost"],
                         specialize=e.get("specialize", False), io_lane_bits=e.get("io_lane_bits", 6),
                         rb=e.get("rb", 4), lds_split=e.get("lds_split", 0), remap=e.get("remap", False))
    return replace(cfg, tb=min(cfg.tb, n - cfg.rb))


def tune_adjoint(n, layers, dtype, device, space=None, repeats=5, log=print, out_path=None):
    """Time energy+gradient of the n-qubit hardware-efficient ansatz for each
    tile configuration in ``space``; cache the fastest."""
    import torch

    from .adjoint import AdjointModel
    from .workloads import hea_circuit, hea_theta, tfim_terms
    space = space or dict(tb=[4, 5, 6, 7, 8], max_cost=[48.0, 1e9], specialize=[False, True], io_lane_bits=[6])
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    th = torch.tensor(theta, device=device)
    results = []
    for tb, mc, spz, io, rb, ls, rm in itertools.product(space["tb"], space["max_cost"], space["specialize"],
                                                         space.get("io_lane_bits", [6]), space.get("rb", [4]),
                                                         space.get("lds_split", [0]), space.get("remap", [False])):
        if tb + rb > n or (rb != 4 and not spz) or (ls and (not spz or ls >= rb)) or (rm and not spz):
            continue
        cfg = TileConfig(tb=tb, low_bits=4, max_cost=mc, specialize=spz, io_lane_bits=io, rb=rb, lds_split=ls,
                         remap=rm)
        m = AdjointModel(circ, tfim_terms(n), len(theta), device=device, dtype=dtype, cfg=cfg)
        for _ in range(2):
            m.energy_and_grad(th)
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for _ in range(repeats):
            m.energy_and_grad(th)
        torch.cuda.synchronize(device)
        t = (time.perf_counter() - t0) / repeats
        results.append((t, cfg))
        log(f"[tune-adjoint] n={n} tb={tb} rb={rb} cap={mc:g} specialize={spz} io={io} lds_split={ls} remap={rm}: "
            f"{t * 1e3:.3f} ms")
    results.sort(key=lambda x: x[0])
    t, cfg = results[0]
    cache = _load()
    cache[f"adjoint:{str(dtype).replace('torch.', '')}:{n}"] = dict(tb=cfg.tb, low_bits=cfg.low_bits,
                                                                     max_cost=cfg.max_cost,
                                                                     specialize=cfg.specialize,
                                                                     io_lane_bits=cfg.io_lane_bits, rb=cfg.rb,
                                                                     lds_split=cfg.lds_split, remap=cfg.remap,
                                                                     ms=t * 1e3)
    with open(out_path or CACHE, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    return cfg, results


def batch_tuning(dtype):
    """The persistent batch's tuned settings for ``dtype`` (empty if untuned):
    {"tasks_per_step": int, "remap_adjoint": {str(n): bool}, "threads": int}."""
    return _load().get(f"batch:{str(dtype).replace('torch.', '')}", {})


def tune_batch(device, dtype=None, nc=512, seed=None, reps=3, log=print, out_path=None):
    """Tune the persistent batch on the 512-circuit workload: for each qubit
    count, whether adjoint sweeps remap qubits (timed on the sub-batch of the
    circuits of that size), then the task granularity (tasks per step) on the
    whole batch with those choices, then (complex64) the workgroup size,
    512 or 256 threads (tiles of 2^12 or 2^11 amplitudes; 3 register bits,
    the pattern library the generic kernels are compiled with).  Writes
    "batch:<dtype>" to the cache."""
    import numpy as np
    import torch

    from .persistent import BatchPlan, PersistentBatch
    from .workloads import SEED_BASE, batch_specs, hea_ops, tfim_terms
    dtype = dtype or torch.complex64
    seed = SEED_BASE + 4 if seed is None else seed
    ns, ls, thetas, _ = batch_specs(nc, seed)
    ns = [int(x) for x in ns]
    ls = [int(x) for x in ls]

    def timed(idx, remap, tps, threads=512):
        plan = BatchPlan([ns[i] for i in idx], [hea_ops(ns[i], ls[i]) for i in idx],
                         [tfim_terms(ns[i]) for i in idx], [len(thetas[i]) for i in idx], dtype,
                         keys=[(ns[i], ls[i]) for i in idx], remap_adjoint=remap, tasks_per_step=tps,
                         threads=threads)
        b = PersistentBatch(None, None, None, None, device=device, dtype=dtype, plan=plan)
        b.set_theta([thetas[i] for i in idx])
        b.capture()
        b.run()
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        torch.cuda.synchronize(device)
        b.check()
        return (time.perf_counter() - t0) / reps

    remap = {}
    for n in sorted(set(ns)):
        idx = [i for i in range(nc) if ns[i] == n]
        t = {rm: timed(idx, rm, 128) for rm in (True, False)}
        remap[str(n)] = bool(min(t, key=t.get))
        log(f"[tune-batch] n={n} ({len(idx)} circuits): remapped adjoint {t[True] * 1e3:.2f} ms, "
            f"natural {t[False] * 1e3:.2f} ms -> remap={remap[str(n)]}")
    full = list(range(nc))
    tt = {}
    for tps in (64, 128, 256):
        tt[tps] = timed(full, remap, tps)
        log(f"[tune-batch] {nc} circuits, {tps} tasks per step: {tt[tps] * 1e3:.1f} ms")
    best = min(tt, key=tt.get)
    threads, ms = 512, tt[best]
    if dtype == torch.complex64:
        t256 = timed(full, remap, best, 256)
        log(f"[tune-batch] {nc} circuits, {best} tasks per step, 256-thread workgroups: {t256 * 1e3:.1f} ms")
        if t256 < ms:
            threads, ms = 256, t256
    cache = _load()
    cache[f"batch:{str(dtype).replace('torch.', '')}"] = dict(tasks_per_step=best, remap_adjoint=remap,
                                                               threads=threads, ms=round(ms * 1e3, 2))
    with open(out_path or CACHE, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    return best, remap
