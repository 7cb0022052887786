"""Ahead-of-time compilation of the per-pass kernels the library's standard
workloads use, so that a fresh install runs them without waiting for hipcc.

The per-pass compiled kernels (qaravan_amd.specialize) are keyed by the text
of the pass structure, not by angles or matrices, so compiling a workload's
program once covers every run of that structure.  ``build()`` calls
``precompile()`` after building the library; everything else (the
generic kernels, the persistent batch) is already inside libqaravan_hip.so.
Workloads covered:

* the 28-qubit, 40-layer brickwork circuit through ``StatevectorSim``'s
  default (tuned) configuration, complex64, with the two variants the
  full-size parity tests run (the circuit then its inverse, its first 6
  layers);
* the 20-qubit, 16-layer hardware-efficient ansatz gradient (forward and
  adjoint programs of the tuned adjoint configuration), complex64;
* for the ``StatevectorSim`` programs, also the variant whose first pass
  generates |0…0⟩ (what ``run()`` executes from the default initial state);
* the reference's published GPU workload, two_local_circ(brickwall_skeleton(n, 1))
  at n = 20 ... 26, where ``StatevectorSim`` switches to the tuned kernels.
"""
import time

import torch

from . import specialize as SP


def _programs():
    from .adjoint import ParamProgram
    from .autotune import best_adjoint_config, sim_tile_config
    from .circuits import two_local_circ
    from .program import compile_ops, lower_circuit
    from .skeletons import brickwall_skeleton
    from .workloads import brickwork_circuit, hea_circuit, hea_theta, tfim_terms
    c64 = torch.complex64
    out = []
    n = 28
    prog = compile_ops(lower_circuit(brickwork_circuit(n, 40).build(), n), n, sim_tile_config(n, c64))
    out.append(("brickwork 28q x 40, complex64", prog, False))
    # the full-size parity tests: the same circuit followed by its inverse, and its first 6 layers
    from .circuits import Circuit
    circ = brickwork_circuit(n, 40)
    inv = Circuit(circ.gate_list + [g.dag() for g in circ.gate_list[::-1]], n=n)
    out.append(("brickwork 28q x 40 then its inverse, complex64",
                compile_ops(lower_circuit(inv.build(), n), n, sim_tile_config(n, c64)), False))
    out.append(("brickwork 28q x 6, complex64",
                compile_ops(lower_circuit(brickwork_circuit(n, 6).build(), n), n, sim_tile_config(n, c64)), False))
    n = 20
    # the program AdjointModel builds from the circuit (the benchmark's path)
    pp = ParamProgram(hea_circuit(n, 16, hea_theta(n, 16)), tfim_terms(n), 2 * n * 16,
                      cfg=best_adjoint_config(n, c64), dtype=c64, start_zero=True)
    out.append(("HEA 20q x 16 forward", pp.fwd0 or pp.fwd, False))
    out.append(("HEA 20q x 16 adjoint", pp.adj, True))
    for n in (20, 22, 24, 26):
        circ = two_local_circ(brickwall_skeleton(n, 1)).build()
        prog = compile_ops(lower_circuit(circ, n), n, sim_tile_config(n, c64))
        out.append((f"two-local brickwall {n}q", prog, False))
    # StatevectorSim.run() from |0…0⟩ runs the variant whose first pass
    # generates the state (only that pass's kernel differs)
    from .program import zero_start_variant
    for name, prog, dual in list(out):
        z = None if dual or name.startswith("HEA") else zero_start_variant(prog, 0)
        if z is not None:
            out.append((name + ", from |0...0>", z, False))
    return out


def precompile(log=print):
    """Compile (or find in the cache) every listed program's kernels;
    returns {workload: seconds}."""
    done = {}
    for name, prog, dual in _programs():
        if not prog.specialize:
            continue
        t0 = time.time()
        SP.compile_program(prog.blob, 0, dual=dual)
        done[name] = round(time.time() - t0, 1)
        log(f"[precompile] {name}: {done[name]} s")
    return done
