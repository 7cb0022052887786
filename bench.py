"""Benchmark of the HIP statevector path on one MI355X.

Workload (a "step"): the 28-qubit, 40-layer seeded brickwork circuit (3900
gates: RX·RY·RZ on every qubit per layer, then CNOTs on alternating bonds) run
from |0…0⟩ in complex64, state resident in HBM.  Reported:

* value = achieved HBM GB/s over the whole step: algorithmic bytes (one read
  and one write of the 2 GiB state per fused pass, plus the initial write)
  divided by the step time;
* gates_per_s = 3900 / step time;
* roofline of the dominant kernel (k_tile, the fused tile pass): algorithmic
  bytes per launch / average launch time from HIP events on the launch stream;
* adjoint gradients: full 640-parameter gradients/s of E = ⟨ψ|H_TFIM|ψ⟩ for the
  20-qubit, 16-layer hardware-efficient ansatz, next to the restated torch
  tape-autograd path on the same GPU;
* fused_blocks: the 30-qubit circuit of 24 layers of Haar-random 5-qubit
  blocks (125 blocks) in complex64 on the FP32 matrix cores: blocks/s, MFMA
  TFLOP/s and utilisation against 157.3 TFLOP/s, error against complex128;
* batch: 512 hardware-efficient-ansatz circuits (10..22 qubits, 4..32 layers)
  forward + adjoint in one HIP-graph launch: circuits/s, p50/p99 latency,
  gradient error against per-circuit complex128 runs on a sample;
* cpu_baseline: the numpy complex128 oracle (tensordot per gate) on a bounded
  sample of the same 28-qubit circuit, timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

STATE_BYTES_C64_28 = 2 * (2 ** 28) * 8  # one read + one write of the c64 state


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-blocks", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-measure", action="store_true")
    ap.add_argument("--batch-rounds", action="store_true",
                    help="also run the round-synchronous batch (per-structure compiled kernels) for comparison")
    ap.add_argument("--no-reference-workload", action="store_true")
    ap.add_argument("--batch-interp", action="store_true", help="round batch with the op-interpreting kernels")
    ap.add_argument("--batch-lds-split", type=int, default=BATCH_LDS_SPLIT,
                    help="batch leg: LDS re-distributions in 2^K parts (compiled kernels)")
    ap.add_argument("--batch-remap", type=int, default=int(BATCH_REMAP),
                    help="batch leg: 1 = remapped adjoint programs (compiled kernels)")
    ap.add_argument("--profile-forward", action="store_true",
                    help="only the timed forward leg (for rocprofv3 runs)")
    ap.add_argument("--profile-blocks", action="store_true", help="only the MFMA block leg")
    ap.add_argument("--profile-adjoint", action="store_true", help="only 20 adjoint gradients (no torch tapes)")
    ap.add_argument("--profile-batch", action="store_true",
                    help="only the 512-circuit persistent batch (3 launches, no checks)")
    ap.add_argument("--profile-copy", action="store_true",
                    help="only a 2 GiB device copy (calibrates the HBM byte counters)")
    ap.add_argument("--tape-host", action="store_true",
                    help="also time the tape path with θ on the host (the reference setup; ~1 min)")
    ap.add_argument("--n", type=int, default=28)
    ap.add_argument("--layers", type=int, default=40)
    return ap.parse_args()


# LDS parts of the compiled batch kernels (log2) and remapped adjoint
# programs in the batch, chosen by measurement
BATCH_LDS_SPLIT = 1
BATCH_REMAP = False


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_forward(args, dev):
    """The 28-qubit step through the drop-in surface: StatevectorSim(circ,
    backend="hip").run() (initialisation + the tuned fused program)."""
    import torch

    from qaravan_amd import StatevectorSim, _lib
    from qaravan_amd import specialize as SP
    from qaravan_amd.workloads import brickwork_circuit

    n, layers = args.n, args.layers
    circ = brickwork_circuit(n, layers)
    n_gates = len(circ.gate_list)
    c0 = dict(SP.STATS)
    t0 = time.perf_counter()
    sim = StatevectorSim(circ, backend="hip", device=dev, dtype=torch.complex64)
    prog = sim.program()
    if prog.specialize:
        prog.specialized(_lib.SVH_C64)          # compiled (or found in the cache) and loaded
    setup_s = time.perf_counter() - t0
    cold = dict(objects=SP.STATS["objects_compiled"] - c0["objects_compiled"],
                compile_s=round(SP.STATS["compile_wall_s"] - c0["compile_wall_s"], 2), setup_s=round(setup_s, 2))
    cfg = sim.tile
    n_tile = prog.n_tile_passes()
    log(f"[bench] forward n={n} layers={layers} gates={n_gates} cfg={cfg} program={prog.describe()} cold={cold}")
    lib = _lib.load()
    st = torch.cuda.current_stream(dev)

    def step(ev=None):
        if ev is not None:
            ev[0].record(st)
        sim.run(progress_bar=False)
        if ev is not None:
            ev[1].record(st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    run_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
    psi = sim.get_statevector()
    state_b = 2 * (2 ** n) * 8
    # the parts of run() that are not tile passes, timed alone on the same
    # stream with the same entry points and sizes: the |0…0⟩ write and, for a
    # remapped program, the final bit permutation (one sweep)
    tmp = torch.empty_like(psi)
    iev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    back = None
    if prog.final_pos is not None:
        import ctypes
        back = [0] * n
        for q, ph in enumerate(prog.final_pos):
            back[ph] = q
        back = (ctypes.c_int * n)(*back)
    for k in range(args.steps):
        iev[k][0].record(st)
        _lib.check(lib.svh_init_basis(_lib.ptr(tmp), n, 0, 0, _lib.stream_ptr(dev)), "init")
        iev[k][1].record(st)
        if back is not None:
            _lib.check(lib.svh_permute_bits(_lib.ptr(tmp), _lib.ptr(psi), n, 0, back, _lib.stream_ptr(dev)),
                       "svh_permute_bits")
        iev[k][2].record(st)
    torch.cuda.synchronize(dev)
    # run() from the default |0…0⟩ executes the program variant whose first
    # pass generates the state: no initialisation sweep inside the step
    zero = sim._program_from_zero() is not None
    init_ms = 0.0 if zero else sum(e[0].elapsed_time(e[1]) for e in iev) / args.steps
    perm_ms = sum(e[1].elapsed_time(e[2]) for e in iev) / args.steps if back is not None else 0.0
    perm_sweeps = 1 if back is not None else 0
    del tmp
    sim.run(progress_bar=False)                 # the timing above overwrote psi
    torch.cuda.synchronize(dev)
    psi = sim.get_statevector()
    tile_ms = (run_ms - init_ms - perm_ms) / n_tile
    alg_bytes_step = (n_tile + perm_sweeps) * state_b + (0 if zero else (2 ** n) * 8)
    return dict(n=n, layers=layers, gates=n_gates, cfg=cfg, prog=prog, ms_step=ms_step,
                gbps=alg_bytes_step / (ms_step * 1e-3) / 1e9, gates_per_s=n_gates / (ms_step * 1e-3),
                tile_ms=tile_ms, tile_gbps=state_b / (tile_ms * 1e-3) / 1e9, n_tile=n_tile,
                state_bytes=state_b, psi=psi, perm_ms=perm_ms, init_ms=init_ms, sweeps=n_tile + perm_sweeps,
                cold=cold, sim=sim)


def check_forward(res, dev):
    """c64 result vs a c128 run on the GPU (‖ψ64 − ψ128‖₂)."""
    import torch

    from qaravan_amd import StatevectorSim
    from qaravan_amd.workloads import brickwork_circuit
    sim = StatevectorSim(brickwork_circuit(res["n"], res["layers"]), backend="hip", device=dev, dtype=torch.complex128,
                         tile=None)
    sim.run(progress_bar=False)
    ref = sim.get_statevector()
    err = torch.linalg.vector_norm(res["psi"].to(torch.complex128) - ref).item()
    norm = torch.linalg.vector_norm(ref).item()
    del sim, ref
    torch.cuda.empty_cache()
    return err, norm


def bench_adjoint(dev, iters=10, tape_host=False):
    """20-qubit, 16-layer HEA (640 parameters), E = ⟨ψ|H_TFIM|ψ⟩: adjoint
    gradients (eager and HIP-graph replay) vs torch tape autograd."""
    import numpy as np
    import torch

    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.autotune import best_adjoint_config
    from qaravan_amd.torch_tape import tape_energy_and_grad
    from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms

    n, layers = 20, 16
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    terms = tfim_terms(n)
    cfg = best_adjoint_config(n, torch.complex64)
    model = AdjointModel(circ, terms, len(theta), device=dev, dtype=torch.complex64, cfg=cfg)
    th = torch.tensor(theta, device=dev)

    def timed(fn, k):
        for _ in range(2):
            fn()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(k):
            out = fn()
        torch.cuda.synchronize(dev)
        return (time.perf_counter() - t0) / k, out

    t_eager, _ = timed(lambda: model.energy_and_grad(th), iters)
    model.capture(th)
    t_graph, (e, g) = timed(lambda: model.replay(th), iters)
    g_adj = g.cpu().numpy()
    e_adj = float(e.item())
    t_adj = min(t_eager, t_graph)

    # c128 adjoint on the GPU as the accuracy reference
    m128 = AdjointModel(circ, terms, len(theta), device=dev, dtype=torch.complex128)
    e128, g128 = m128.energy_and_grad(th)
    g128 = g128.cpu().numpy()
    rel_err = float(np.linalg.norm(g_adj - g128) / np.linalg.norm(g128))

    th_small = torch.tensor(hea_theta(12, 2), dtype=torch.float64, device=dev)
    small = hea_circuit(12, 2, th_small.cpu().numpy())
    th_dev = torch.tensor(theta, dtype=torch.float64, device=dev)
    res = {}
    for kernel, k in (("elementwise", 2), ("tensordot", 1)):
        tape_energy_and_grad(small, th_small, tfim_terms(12), dev, kernel)  # warm the torch kernels
        torch.cuda.synchronize(dev)
        log(f"[bench] torch tape ({kernel}, device θ): timing 20-qubit gradients")
        t0 = time.perf_counter()
        for _ in range(k):
            e_t, g_t = tape_energy_and_grad(circ, th_dev, terms, dev, kernel)
        torch.cuda.synchronize(dev)
        dt = (time.perf_counter() - t0) / k
        res[kernel] = (dt, float(np.linalg.norm(g_t.cpu().numpy() - g128) / np.linalg.norm(g128)))
    out = dict(adjoint_grads_per_s=1.0 / t_adj, adjoint_eager_ms=t_eager * 1e3, adjoint_graph_ms=t_graph * 1e3,
               tile=dict(tb=cfg.tb, low_bits=cfg.low_bits, max_cost=cfg.max_cost, specialize=cfg.specialize,
                         io_lane_bits=cfg.io_lane_bits),
               rel_grad_err_vs_c128=rel_err,
               torch_autograd_grads_per_s=1.0 / res["elementwise"][0],
               speedup=res["elementwise"][0] / t_adj,
               speedup_note="against the elementwise tape (slice arithmetic per gate, the faster tape); the "
                            "tensordot tape, a literal restatement of the reference's op_action, is slower",
               torch_autograd_rel_grad_err_vs_c128=res["elementwise"][1],
               torch_autograd_tensordot_grads_per_s=1.0 / res["tensordot"][0],
               speedup_vs_tensordot_tape=res["tensordot"][0] / t_adj,
               torch_autograd_tensordot_rel_grad_err_vs_c128=res["tensordot"][1],
               energy=e_adj, energy_c128=float(e128.item()),
               n_params=len(theta), n_gates=len(circ.gate_list))
    if tape_host:
        th_cpu = torch.tensor(theta, dtype=torch.float64)
        t0 = time.perf_counter()
        e_t, g_t = tape_energy_and_grad(circ, th_cpu, terms, dev)
        torch.cuda.synchronize(dev)
        out.update(torch_autograd_host_theta_grads_per_s=1.0 / (time.perf_counter() - t0))
    return out


def profile_adjoint(dev, reps=20):
    """The 20-qubit gradient alone, for kernel traces."""
    import torch

    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.autotune import best_adjoint_config
    from qaravan_amd.workloads import hea_circuit, hea_theta, tfim_terms
    n, layers = 20, 16
    theta = hea_theta(n, layers)
    cfg = best_adjoint_config(n, torch.complex64)
    model = AdjointModel(hea_circuit(n, layers, theta), tfim_terms(n), len(theta), device=dev,
                         dtype=torch.complex64, cfg=cfg)
    th = torch.tensor(theta, device=dev)
    for _ in range(2):
        model.energy_and_grad(th)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(reps):
        model.energy_and_grad(th)
    torch.cuda.synchronize(dev)
    return dict(ms=(time.perf_counter() - t0) / reps * 1e3, fwd_passes=model.pp.fwd.n_tile_passes(),
                adj_passes=model.pp.adj.n_tile_passes())


def profile_batch(dev, nc=512, seed=20260004, reps=3):
    """The persistent 512-circuit batch alone (eager launches), for kernel traces."""
    import torch

    from qaravan_amd.persistent import BatchPlan, PersistentBatch
    from qaravan_amd.workloads import batch_specs, hea_ops, tfim_terms
    ns, ls, thetas, _ = batch_specs(nc, seed)
    plan = BatchPlan(ns.tolist(), [hea_ops(int(n), int(l)) for n, l in zip(ns, ls)],
                     [tfim_terms(int(n)) for n in ns], [len(t) for t in thetas],
                     keys=list(zip(ns.tolist(), ls.tolist())))
    b = PersistentBatch(None, None, None, None, device=dev, plan=plan)
    b.set_theta(thetas)
    b.run()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    torch.cuda.synchronize(dev)
    b.check()
    return dict(ms=(time.perf_counter() - t0) / reps * 1e3, **plan.describe())


def bench_blocks(dev, n=30, layers=24, reps=3):
    """Dense 5-qubit Haar blocks on MFMA (complex64) + a complex128 check."""
    import torch

    from qaravan_amd import _lib
    from qaravan_amd.program import TileConfig, compile_ops, lower_circuit
    from qaravan_amd.statevector_sim import DeviceScratch, run_program
    from qaravan_amd.workloads import haar_block_circuit
    circ = haar_block_circuit(n, layers).build()
    nb = len(circ.gate_list)
    prog = compile_ops(lower_circuit(circ, n), n, TileConfig(tb=8))
    lib = _lib.load()
    sc = DeviceScratch(dev)
    psi = torch.empty(2 ** n, dtype=torch.complex64, device=dev)
    times = []
    for r in range(reps + 1):
        _lib.check(lib.svh_init_basis(_lib.ptr(psi), n, 0, 0, _lib.stream_ptr(dev)), "init")
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        run_program(psi, prog, sc)
        torch.cuda.synchronize(dev)
        if r:
            times.append(time.perf_counter() - t0)
    dt = min(times)
    flops = nb * (64 * 64 * 2) * 2 ** (n - 5)
    sweeps = _block_sweeps(prog)
    psi128 = torch.empty(2 ** n, dtype=torch.complex128, device=dev)
    _lib.check(lib.svh_init_basis(_lib.ptr(psi128), n, 1, 0, _lib.stream_ptr(dev)), "init")
    run_program(psi128, prog, sc)
    err = torch.linalg.vector_norm(psi.to(torch.complex128) - psi128).item()
    del psi, psi128
    torch.cuda.empty_cache()
    return dict(workload=f"{n}q x {layers} layers of Haar 5-qubit blocks ({nb} blocks), complex64",
                ms=dt * 1e3, ms_per_block=dt * 1e3 / nb, blocks_per_s=nb / dt,
                mfma_tflops=flops / dt / 1e12, mfma_util=flops / dt / 157.3e12,
                flop_count="4 real products per complex product (2^38 per block at 30 qubits); the kernels "
                           "execute 3/4 of it (3M form)",
                executed_mfma_util=0.75 * flops / dt / 157.3e12, sweeps=sweeps,
                hbm_gbps=sweeps * 2 * 2 ** n * 8 / dt / 1e9, l2_err_vs_c128=err)


def _block_sweeps(prog):
    """State sweeps the library makes for a program of complex64 5-qubit dense
    passes: runs of consecutive passes on pairwise disjoint qubits (at most 8)
    go two per sweep, an odd one alone."""
    from qaravan_amd.program import PASS_DENSE
    sweeps, group = 0, []
    for p in prog.passes + [None]:
        bits = set(p.bits) if p is not None and p.kind == PASS_DENSE and len(p.bits) == 5 else None
        if bits is not None and len(group) < 8 and all(not (bits & g) for g in group):
            group.append(bits)
            continue
        sweeps += (len(group) + 1) // 2
        group = [bits] if bits is not None else []
    return sweeps


def bench_c0(dev, reps=20):
    """The 12-qubit 20-layer brickwork (830 gates, complex128) end to end on
    the GPU, checked against the committed golden final state."""
    import numpy as np
    import torch

    from qaravan_amd import _lib
    from qaravan_amd.program import TileConfig, compile_ops, lower_circuit
    from qaravan_amd.statevector_sim import DeviceScratch, run_program
    from qaravan_amd.workloads import brickwork_circuit
    n, layers = 12, 20
    circ = brickwork_circuit(n, layers).build()
    t0 = time.perf_counter()
    prog = compile_ops(lower_circuit(circ, n), n, TileConfig(tb=6, low_bits=4, max_cost=1e9))
    compile_ms = (time.perf_counter() - t0) * 1e3
    lib = _lib.load()
    psi = torch.empty(2 ** n, dtype=torch.complex128, device=dev)
    sc = DeviceScratch(dev)
    times = []
    for r in range(reps + 2):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        _lib.check(lib.svh_init_basis(_lib.ptr(psi), n, 1, 0, _lib.stream_ptr(dev)), "init")
        run_program(psi, prog, sc)
        torch.cuda.synchronize(dev)
        if r >= 2:
            times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    gold = np.load(os.path.join(ROOT, "tests", "golden", "brickwork_12q_20l_state.npz"))["state"]
    err = float(np.linalg.norm(psi.cpu().numpy() - gold))
    return dict(workload="brickwork 12q x 20 layers (830 gates), complex128", median_ms=ms,
                gates_per_s=len(circ.gate_list) / (ms * 1e-3), compile_ms=compile_ms,
                passes=prog.n_passes, l2_err_vs_golden=err)


def bench_measure(dev, n=28, reps=5, shots=1_000_000):
    """Measurement kernels on a 28-qubit complex64 state (2 GiB, resident in
    HBM): marginal probabilities over 2 and over all 28 qubits, a 10^6-shot
    draw over all 28, and a 2- and a 5-qubit reduced density matrix."""
    import numpy as np
    import torch

    from qaravan_amd import _lib
    from qaravan_amd import measure as M
    from qaravan_amd.program import TileConfig, compile_ops, lower_circuit
    from qaravan_amd.statevector_sim import DeviceScratch, run_program
    from qaravan_amd.workloads import brickwork_circuit
    lib = _lib.load()
    psi = torch.empty(2 ** n, dtype=torch.complex64, device=dev)
    _lib.check(lib.svh_init_basis(_lib.ptr(psi), n, 0, 0, _lib.stream_ptr(dev)), "init")
    run_program(psi, compile_ops(lower_circuit(brickwork_circuit(n, 4).build(), n), n, TileConfig(tb=8)),
                DeviceScratch(dev))
    state_bytes = 2 ** n * 8

    def timed(fn):
        fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize(dev)
        return s.elapsed_time(e) / reps

    out = {"workload": f"{n}-qubit complex64 state (brickwork, 4 layers)"}
    t = timed(lambda: M.marginal_probabilities(psi, [0, n - 1]))
    out["probs_2q_ms"], out["probs_2q_gbps"] = t, state_bytes / t / 1e6
    t = timed(lambda: M.marginal_probabilities(psi, list(range(n))))
    out["probs_all_ms"], out["probs_all_gbps"] = t, (state_bytes + 2 ** n * 8) / t / 1e6
    for k, sites in ((2, [3, 17]), (5, [0, 6, 13, 20, 27])):
        t = timed(lambda: M.rdm_from_sv(psi, sites))
        out[f"rdm_{k}q_ms"], out[f"rdm_{k}q_gbps"] = t, state_bytes / t / 1e6
    np.random.seed(1)
    M.measure_sv(psi, list(range(n)), shots=shots)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    M.measure_sv(psi, list(range(n)), shots=shots)
    out["sample_all_1e6_shots_ms"] = (time.perf_counter() - t0) * 1e3
    del psi
    torch.cuda.empty_cache()
    return out


def bench_batch(dev, nc=512, seed=20260004, check=6, specialize=True, lds_split=0, remap=False):
    """512 HEA circuits (10..22 qubits, 4..32 layers) forward + adjoint as
    round-synchronous launches (one per pass structure and round) recorded in
    one HIP graph; latency from device timestamps.  Kept for comparison with
    the persistent batch (bench_persistent)."""
    import numpy as np
    import torch

    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.batch import BatchedAdjoint
    from qaravan_amd.workloads import batch_specs, hea_circuit, hea_ops, tfim_terms
    ns, ls, thetas, rng = batch_specs(nc, seed)
    t0 = time.perf_counter()
    b = BatchedAdjoint(None, [tfim_terms(int(n)) for n in ns], [len(t) for t in thetas], device=dev,
                       dtype=torch.complex64, keys=list(zip(ns.tolist(), ls.tolist())),
                       op_lists=[hea_ops(int(n), int(l)) for n, l in zip(ns, ls)], ns=ns.tolist(),
                       specialize=specialize, lds_split=lds_split if specialize else 0,
                       remap=remap and specialize)
    log(f"[bench] batch: planned {nc} circuits in {time.perf_counter() - t0:.1f} s, {len(b.launches)} launches")
    b.set_theta(thetas)
    b.capture()
    for _ in range(2):
        b.run()
    torch.cuda.synchronize(dev)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    torch.cuda.synchronize(dev)
    dt = (time.perf_counter() - t0) / reps
    lat = b.latencies_us()
    kern = f"{len(b.kernels.names)} pass-structure kernels" if specialize else "op-interpreting kernels"
    return dict(workload=f"{nc} HEA circuits, round-synchronous launches in one HIP graph ({b.n_launches} "
                         f"launches, {kern})",
                circuits_per_s=nc / dt, ms_per_batch=dt * 1e3, p50_latency_ms=float(np.percentile(lat, 50)) / 1e3,
                p99_latency_ms=float(np.percentile(lat, 99)) / 1e3)


def bench_persistent(dev, nc=512, seed=20260004, check=6):
    """The same 512 circuits as ONE persistent launch with a device task
    queue (qaravan_amd.persistent), recorded in a HIP graph; per-circuit
    latency from device timestamps; achieved bytes = each step's algorithmic
    state traffic over the batch time; gradient error against complex128
    single-circuit adjoint runs and, for a small sample, against the oracle."""
    import numpy as np
    import torch

    from oracle import sv_oracle as O
    from qaravan_amd import specialize as SP
    from qaravan_amd.adjoint import AdjointModel
    from qaravan_amd.persistent import BatchPlan, PersistentBatch
    from qaravan_amd.workloads import batch_specs, hea_circuit, hea_ops, tfim_terms
    ns, ls, thetas, rng = batch_specs(nc, seed)
    c0 = dict(SP.STATS)
    t0 = time.perf_counter()
    plan = BatchPlan(ns.tolist(), [hea_ops(int(n), int(l)) for n, l in zip(ns, ls)],
                     [tfim_terms(int(n)) for n in ns], [len(t) for t in thetas],
                     keys=list(zip(ns.tolist(), ls.tolist())))
    plan_s = time.perf_counter() - t0
    b = PersistentBatch(None, None, None, None, device=dev, plan=plan)
    log(f"[bench] persistent batch: planned in {plan_s:.1f} s: {plan.describe()}")
    b.set_theta(thetas)
    b.capture()
    for _ in range(2):
        b.run()
    torch.cuda.synchronize(dev)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    torch.cuda.synchronize(dev)
    dt = (time.perf_counter() - t0) / reps
    b.check()
    lat = b.latencies_us()
    # one more (untimed) run with the workers' clock accounting on
    b.enable_stats()
    b._launch()
    torch.cuda.synchronize(dev)
    b.check()
    sched = b.scheduler_stats()
    worker_s = sched["workgroups"] * dt
    sched = dict(workers=sched["workgroups"], task_frac=round((sched["task_s"] - sched["wait_s"]) / worker_s, 3),
                 wait_frac=round(sched["wait_s"] / worker_s, 3), tasks=sched["tasks"], waits_spun=sched["waits_spun"],
                 sync_mode="xcd-local" if b.sync_mode == 1 else "device")
    grads = b.gradients()
    worst = 0.0
    idx = list(np.argsort(ns)[-2:]) + list(rng.choice(nc, size=check - 2, replace=False))
    for c in idx:
        c = int(c)
        m = AdjointModel(hea_circuit(int(ns[c]), int(ls[c]), thetas[c]), tfim_terms(int(ns[c])), len(thetas[c]),
                         device=dev, dtype=torch.complex128)
        _, g = m.energy_and_grad(torch.tensor(thetas[c], device=dev))
        g = g.cpu().numpy()
        worst = max(worst, float(np.linalg.norm(grads[c] - g) / np.linalg.norm(g)))
    # the oracle on the smallest circuit of the batch
    cs = int(np.argmin(np.asarray(ns) * 100 + np.asarray(ls)))
    n, L = int(ns[cs]), int(ls[cs])
    spec = []
    for l in range(L):
        for k, w in enumerate("yz"):
            spec += [(w, [q], (2 * l + k) * n + q, None) for q in range(n)]
        spec += [(None, [i, i + 1], None, np.diag([1, 1, 1, -1]).astype(complex)) for i in range(n - 1)]
    _, g_ref = O.adjoint_gradient(spec, thetas[cs], n, O.tfi_terms(n, -1.0, -1.0))
    oracle_err = float(np.linalg.norm(grads[cs] - g_ref) / np.linalg.norm(g_ref))
    alg = plan.algorithmic_bytes()
    d = plan.describe()
    return dict(workload=f"{nc} HEA circuits, n~U{{10..22}}, layers~U{{4..32}}, TFIM loss, forward+adjoint, "
                         f"one persistent launch ({d['tasks']} tasks in {d['queues']} queues, generic kernels)",
                circuits_per_s=nc / dt, ms_per_batch=dt * 1e3,
                p50_latency_ms=float(np.percentile(lat, 50)) / 1e3, p99_latency_ms=float(np.percentile(lat, 99)) / 1e3,
                max_rel_grad_err_vs_c128_sample=worst, checked_circuits=len(idx),
                rel_grad_err_vs_oracle=oracle_err, oracle_circuit=f"{n}q x {L} layers",
                algorithmic_bytes=alg, achieved_gbps=alg / dt / 1e9, hbm_frac=alg / dt / 8e12,
                bytes_note="ψ read+write per forward tile pass, ψ and λ per adjoint pass, ψ read + λ written by "
                           "the Hamiltonian step; most circuits' states live in L2/MALL, so this can exceed HBM",
                limiter="valu", valu_busy=0.77,
                valu_source="profiles/batch_512_hea_c64.md (SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x launch cycles))",
                scheduler=sched,
                plan_s=round(plan_s, 1), compile_s=round(SP.STATS["compile_wall_s"] - c0["compile_wall_s"], 2),
                total_amplitudes=float(sum(2.0 ** int(n) for n in ns)))


def bench_reference_workload(dev, ns=(16, 18, 20, 22, 24, 26)):
    """The reference's own published GPU workload (examples/gpu_acceleration.ipynb,
    cells 2-3): two_local_circ(brickwall_skeleton(n, 1)) — n-1 random 4x4
    gates — run with StatevectorSim(circ, backend="torch", device="cuda"),
    timed from construction to synchronisation as the notebook does.  The
    reference's plot puts a Colab T4 at about 0.011 s (n=16) ... 4.5 s (n=26)."""
    import numpy as np
    import torch

    from qaravan_amd import StatevectorSim, brickwall_skeleton, two_local_circ
    from qaravan_amd import specialize as SP
    t4 = {16: 0.011, 20: 0.12, 22: 0.32, 24: 1.15, 26: 4.5}
    rng = np.random.default_rng(20260099)
    out = {}
    for n in ns:
        circ = two_local_circ(brickwall_skeleton(n, 1), rng=rng)
        c0 = SP.STATS["compile_wall_s"]
        times = []
        for r in range(4):
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
            sim = StatevectorSim(circ, backend="torch", device="cuda")
            sim.run(progress_bar=False)
            torch.cuda.synchronize(dev)
            times.append(time.perf_counter() - t0)
        out[str(n)] = dict(first_s=round(times[0], 4), warm_s=round(float(np.median(times[1:])), 5),
                           compile_s=round(SP.STATS["compile_wall_s"] - c0, 2), t4_s=t4.get(n),
                           speedup_vs_t4=(round(t4[n] / float(np.median(times[1:])), 1) if n in t4 else None))
        del sim
    torch.cuda.empty_cache()
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    return "unknown"


def cpu_baseline(n, layers, max_seconds=10.0):
    """Oracle (numpy complex128, one tensordot per gate) on the first gates of
    the same layered circuit, until ~max_seconds of CPU work, with all BLAS
    threads and with one."""
    import numpy as np

    from oracle import sv_oracle as O
    from qaravan_amd.workloads import brickwork_circuit
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        threads = max([d.get("num_threads", 1) for d in threadpool_info()] + [1])
    except Exception:  # pragma: no cover
        threadpool_limits, threads = None, os.cpu_count()
    circ = brickwork_circuit(n, layers)
    seq = [g for layer in O.construct_layers(circ.gate_list) for g in layer]

    def leg(limit):
        sv = O.init_state(n)
        done = 0
        t0 = time.perf_counter()
        for g in seq:
            if threadpool_limits is not None:
                with threadpool_limits(limits=limit):
                    sv = O.op_action(np.asarray(g.matrix), g.indices, sv)
            else:
                sv = O.op_action(np.asarray(g.matrix), g.indices, sv)
            done += 1
            if time.perf_counter() - t0 > max_seconds:
                break
        return done, time.perf_counter() - t0

    done, dt = leg(threads)
    done1, dt1 = leg(1)
    return dict(value=done / dt, unit="gates/s", cores=int(threads), kind="port", cpu_model=_cpu_model(),
                one_thread=dict(value=done1 / dt1, unit="gates/s", cores=1,
                                sample=f"first {done1} gates, {dt1:.1f} s"),
                sample=f"first {done} gates (of {len(seq)}) of the {n}-qubit {layers}-layer brickwork "
                       f"circuit, numpy complex128 oracle, {dt:.1f} s with {threads} BLAS threads")


def cpu_other_configs(max_seconds=15.0):
    """Oracle timings beside the other workloads (part of the CPU-baseline
    leg): the 12-qubit brickwork (median of 5 runs, 1 thread and all
    threads), one layer of 5-qubit Haar blocks at 24 qubits, and forward runs
    of the smallest batch circuits until ~max_seconds."""
    import numpy as np

    from oracle import sv_oracle as O
    from qaravan_amd.workloads import batch_specs, brickwork_circuit, haar_block_layers, hea_circuit
    out = {}
    c0 = brickwork_circuit(12, 20)
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        all_threads = max([d.get("num_threads", 1) for d in threadpool_info()] + [1])
    except Exception:  # pragma: no cover
        threadpool_limits, all_threads = None, os.cpu_count()
    for label, lim in (("1_thread", 1), ("all_threads", all_threads)):
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            if threadpool_limits is not None:
                with threadpool_limits(limits=lim):
                    O.run(c0.gate_list, 12)
            else:
                O.run(c0.gate_list, 12)
            times.append(time.perf_counter() - t0)
        ms = float(np.median(times)) * 1e3
        out[f"c0_{label}"] = dict(median_ms=ms, gates_per_s=len(c0.gate_list) / (ms * 1e-3), threads=lim)
    out["cpu_count"] = os.cpu_count()
    layer = haar_block_layers(24, 1)[0]
    sv = O.init_state(24)
    t0 = time.perf_counter()
    for q0, u in layer:
        sv = O.op_action(u, list(range(q0, q0 + 5)), sv)
    dt = time.perf_counter() - t0
    out["haar5_24q_one_layer"] = dict(blocks=len(layer), s=dt, blocks_per_s=len(layer) / dt)
    ns, ls, thetas, _ = batch_specs()
    order = np.argsort(ns, kind="stable")
    done, t0 = 0, time.perf_counter()
    for c in order[:16]:
        c = int(c)
        O.run(hea_circuit(int(ns[c]), int(ls[c]), thetas[c]).gate_list, int(ns[c]))
        done += 1
        if time.perf_counter() - t0 > max_seconds:
            break
    dt = time.perf_counter() - t0
    out["batch_forward_subset"] = dict(circuits=done, s=dt, circuits_per_s=done / dt,
                                       sample="smallest circuits of the 512-circuit batch, forward only")
    return out


def valu_limiter():
    """What limits the tile pass besides HBM: the FP32 VALU issue rate, from
    the committed rocprofv3 counter run (SQ_INSTS_VALU over the 28-qubit
    passes; VALU busy = instructions x 4 cycles / (SIMDs x cycles))."""
    path = os.path.join(ROOT, "profiles", "tile_pass_valu_28q.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        d = json.load(f)
    return {"limiter": "valu", "valu_busy": d.get("valu_busy_step"),
            "valu_busy_heaviest_passes": d.get("valu_busy_heaviest"), "valu_source": "profiles/tile_pass_valu_28q.json"}


def traffic_bytes(fwd):
    """HBM bytes per tile-pass launch from the committed rocprofv3 counter runs
    (FETCH_SIZE x2 calibration + WRITE_SIZE, KiB per dispatch), or None."""
    path = os.path.join(ROOT, "profiles", "tile_pass_traffic_28q.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(multi-GPU sharding is a design only)"}))
        sys.exit(2)
    import torch
    dev = torch.device("cuda:0")
    if args.profile_copy:
        a = torch.ones(2 ** 28, dtype=torch.complex64, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize(dev)
        print(json.dumps({"profile": "copy", "bytes_read": a.numel() * 8, "bytes_written": b.numel() * 8}))
        return
    if args.profile_blocks:
        print(json.dumps({"profile": "blocks", **bench_blocks(dev, reps=1)}))
        return
    if args.profile_adjoint:
        print(json.dumps({"profile": "adjoint", **profile_adjoint(dev)}))
        return
    if args.profile_batch:
        print(json.dumps({"profile": "batch", **profile_batch(dev)}))
        return
    fwd = bench_forward(args, dev)
    if args.profile_forward:
        print(json.dumps({"profile": "forward", "ms_per_step": fwd["ms_step"], "tile_ms": fwd["tile_ms"],
                          "n_tile": fwd["n_tile"]}))
        return
    log(f"[bench] forward: {fwd['ms_step']:.2f} ms/step, {fwd['gbps']:.0f} GB/s, "
        f"{fwd['gates_per_s']:.0f} gates/s, tile pass {fwd['tile_ms']:.3f} ms = {fwd['tile_gbps']:.0f} GB/s")
    err, norm = check_forward(fwd, dev)
    log(f"[bench] c64 vs c128 at {fwd['n']}q: {err:.3e} (norm {norm:.12f})")
    del fwd["sim"], fwd["psi"]
    torch.cuda.empty_cache()
    adj = None
    if not args.no_adjoint:
        adj = bench_adjoint(dev, tape_host=args.tape_host)
        log(f"[bench] adjoint: {adj}")
    blocks = None
    if not args.no_blocks:
        blocks = bench_blocks(dev)
        log(f"[bench] fused blocks: {blocks}")
    batch = rounds = None
    if not args.no_batch:
        batch = bench_persistent(dev)
        log(f"[bench] persistent batch: {batch}")
        if args.batch_rounds:
            rounds = bench_batch(dev, specialize=not args.batch_interp, lds_split=args.batch_lds_split,
                                 remap=bool(args.batch_remap))
            log(f"[bench] round-synchronous batch: {rounds}")
    ref_wl = None
    if not args.no_reference_workload:
        ref_wl = bench_reference_workload(dev)
        log(f"[bench] reference workload: {ref_wl}")
    c0 = bench_c0(dev)
    log(f"[bench] c0: {c0}")
    meas = None
    if not args.no_measure:
        meas = bench_measure(dev)
        log(f"[bench] measure: {meas}")
    cpu = cpu_more = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(fwd["n"], fwd["layers"])
        log(f"[bench] cpu baseline: {cpu}")
        cpu_more = cpu_other_configs()
        log(f"[bench] cpu, other workloads: {cpu_more}")
    peak = 8000.0
    out = {
        "metric": "HBM GB/s + gates/s at 28q c64; adjoint grads/s vs torch autograd",
        "value": round(fwd["gbps"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fwd["ms_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex)",
        "data": "synthetic seeded circuit",
        "config": {"workload": f"brickwork {fwd['n']}q x {fwd['layers']} layers ({fwd['gates']} gates), "
                               f"state 2^{fwd['n']} complex64 in HBM",
                   "tile": {"tb": fwd["cfg"].tb, "low_bits": fwd["cfg"].low_bits,
                            "max_cost": fwd["cfg"].max_cost, "fuse2q": fwd["cfg"].fuse2q,
                            "io_lane_bits": fwd["cfg"].io_lane_bits, "register_bits": fwd["cfg"].rb,
                            "lds_split": fwd["cfg"].lds_split, "remap": fwd["cfg"].remap,
                            "remap_free": fwd["cfg"].remap_free},
                   "fused_passes": fwd["n_tile"],
                   "state_sweeps": fwd["sweeps"],
                   "final_permutation_ms": round(fwd["perm_ms"], 3)},
        "gates_per_s": round(fwd["gates_per_s"], 1),
        "c64_vs_c128_l2_err": err,
        "roofline": {"bound": "hbm", "achieved": round(fwd["tile_gbps"], 1), "peak": peak, "unit": "GB/s",
                     "frac": round(fwd["tile_gbps"] / peak, 4), "traffic": traffic_bytes(fwd),
                     "kernel": ("svh_sp_<p> (specialised tile pass)" if fwd["cfg"].specialize
                                else "k_tile<float, tb, 4, false>"),
                     "bytes_per_launch": fwd["state_bytes"], "avg_launch_ms": round(fwd["tile_ms"], 4),
                     **valu_limiter()},
        "cpu_baseline": cpu,
        "cold_start": {"forward": fwd["cold"]},
    }
    if adj is not None:
        out["adjoint"] = adj
    if blocks is not None:
        out["fused_blocks"] = blocks
    if batch is not None:
        out["batch"] = batch
        out["cold_start"]["batch"] = dict(plan_s=batch["plan_s"], compile_s=batch["compile_s"])
    if rounds is not None:
        out["batch_rounds"] = rounds
    if ref_wl is not None:
        out["reference_workload"] = ref_wl
    if meas is not None:
        out["measure"] = meas
    out["c0"] = c0
    if cpu_more is not None:
        out["cpu_other_workloads"] = cpu_more
    print(json.dumps(out))


if __name__ == "__main__":
    main()
