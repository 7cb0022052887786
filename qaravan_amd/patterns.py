"""Stage-pattern library of the generic tile kernels.

A fused pass is a run of stages; a stage applies a short op list to the
registers of one layout (qaravan_amd.program).  The per-pass compiled kernels
(qaravan_amd.specialize) spell every op out with constant register indices,
which is fast but needs hipcc at run time for every new circuit structure.
The generic kernels (csrc/svh_gtile.hip) instead walk the stage records at run
time and, for each stage, branch once to straight-line code for the stage's
op sequence when that sequence is in this library, or interpret the stage op
by op (a run-time switch per op) when it is not.  Both give the same result;
the library only saves the per-op dispatch.

A pattern is the op sequence of a stage without its data: per op (kind,
register bits a and b, Pauli code, gradient flag).  Matrices, angles, address
tables and gradient slots still come from the records, so one pattern serves
every circuit, angle vector and tile whose stage has that sequence.  Op
sequences recur heavily: the stages of the 512-circuit hardware-efficient
ansatz batch (10-22 qubits, 1-32 layers) reduce to a few dozen patterns per
(register bits, direction).

``python -m qaravan_amd.patterns`` regenerates csrc/svh_patterns.h (device
code) and csrc/svh_patterns.json (the key table read here) from the canonical
workloads listed in ``_canonical_programs``; both are build inputs.
"""
import json
import os

import numpy as np

from . import program as P

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "csrc", "svh_patterns.json")
HEADER = os.path.join(HERE, "csrc", "svh_patterns.h")
LIB_RB = (3,)           # register bits that get compiled patterns
DISPATCH_RB = (3, 4)    # register bits the generic kernels are built for

_table = None


def stage_key(op_rows, dual):
    """The pattern key of a stage's op records (structured OP_DT rows)."""
    return tuple((int(o['kind']), int(o['a']), int(o['b']), int(o['pauli']),
                  1 if (dual and int(o['flags']) & P.OPF_GRAD) else 0) for o in op_rows)


def load_table():
    """{(rb, dual): {key: pattern id}} from the committed table (empty if absent)."""
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(TABLE):
            with open(TABLE) as f:
                raw = json.load(f)
            for name, keys in raw.items():
                rb, dual = (int(x) for x in name.split(":"))
                _table[(rb, dual)] = {tuple(tuple(op) for op in k): i for i, k in enumerate(keys)}
    return _table


def pattern_ids(stages, ops, rb, dual):
    """Pattern id (or -1) of every stage record, for a program's stage/op arrays."""
    lib = load_table().get((rb, int(dual)), {})
    out = np.full(len(stages), -1, dtype=np.int32)
    for s, st in enumerate(stages):
        key = stage_key(ops[int(st['op_begin']):int(st['op_end'])], dual)
        out[s] = lib.get(key, -1) if key else -1
    return out


# ---------------------------------------------------------------- generator
def _op_line(k, op, dual):
    kind, a, b, pauli, grad = op
    hi, lo = max(a, b), min(a, b)
    g = "true" if grad else "false"
    if kind == P.OP_MAT1:
        one = f"sp_mat1<R, {a}, RB>"
        return f"    {one}(v, ops[{k}]);" + (f" {one}(w, ops[{k}]);" if dual else "")
    if kind == P.OP_MAT2:
        one = f"sp_mat2<R, {a}, {b}, RB>"
        return f"    {one}(v, ops[{k}]);" + (f" {one}(w, ops[{k}]);" if dual else "")
    if kind == P.OP_CX:
        one = f"ap_cx_t<R, RB, {hi}, {lo}, {'true' if a > b else 'false'}>"
        return f"    {one}(v);" + (f" {one}(w);" if dual else "")
    if kind == P.OP_CZ:
        one = f"ap_cz<R, RB, {hi}, {lo}>"
        return f"    {one}(v);" + (f" {one}(w);" if dual else "")
    if kind == P.OP_ROT1:
        if dual:
            # the gradient term is taken by the caller (see _pattern_lines)
            pre = f"    const double gq{k} = gr_rot1<R, RB, {a}>(v, w, {pauli});\n" if grad else ""
            return pre + f"    gd_rot1<R, RB, {a}, {pauli}, false>(v, w, ops[{k}], trig, gout, lane0);"
        return f"    sp_rot1<R, {a}, {pauli}, RB>(v, ops[{k}], trig);"
    if kind == P.OP_ROT2:
        if dual:
            pre = f"    const double gq{k} = gr_rot2<R, RB, {hi}, {lo}>(v, w, {pauli});\n" if grad else ""
            return pre + f"    gd_rot2<R, RB, {hi}, {lo}, {pauli}, false>(v, w, ops[{k}], trig, gout, lane0);"
        return f"    sp_rot2<R, {a}, {b}, {pauli}, RB>(v, ops[{k}], trig);"
    raise ValueError(kind)


def _pattern_lines(key, dual):
    """Body of one pattern.  In an adjoint stage each gradient term is
    computed where its rotation is undone and kept in a register; the terms
    (consecutive slots) are then summed over the wave together, up to 8 at a
    time, and stored to the wave's partial row."""
    lines, grads = [], []

    def flush():
        if grads:
            cnt = len(grads)
            K = (cnt - 1).bit_length()
            vals = ", ".join([f"gq{k}" for k in grads] + ["0.0"] * ((1 << K) - cnt))
            lines.append(f"    {{ double gv[{1 << K}] = {{{vals}}}; "
                         f"gd_grad_group<{K}>(gv, gout, ops[{grads[0]}].grad_slot, {cnt}); }}")
            grads.clear()
    for k, op in enumerate(key):
        lines.append(_op_line(k, op, dual))
        if dual and op[4] and op[0] in (P.OP_ROT1, P.OP_ROT2):
            grads.append(k)
            if len(grads) == 8:
                flush()
    flush()
    return lines


def generate(keys_by_lib):
    """(header text, json text) for {(rb, dual): [keys]}."""
    out = ["// Generated by `python -m qaravan_amd.patterns` (do not edit): straight-line",
           "// code for the stage op sequences of the pattern library (svh_patterns.json).",
           "#pragma once", '#include "svh_gtile.h"', "", "namespace svh {", ""]
    for (rb, dual), keys in sorted(keys_by_lib.items()):
        tag = f"{rb}{'d' if dual else 'f'}"
        args = ("cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, "
                "const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0") if dual else \
               "cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig"
        for i, key in enumerate(keys):
            out.append(f"template <typename R, int RB>\n__device__ __forceinline__ void pat{tag}_{i}({args}) {{")
            out += _pattern_lines(key, dual)
            out.append("}")
        call = "v, w, ops, trig, gout, lane0" if dual else "v, ops, trig"
        out.append(f"// pattern {tag}: returns false when pid is not in the library")
        out.append(f"template <typename R, int RB>\n__device__ __forceinline__ bool run_pattern{tag}(int pid, {args}) {{")
        out.append("    switch (pid) {")
        out += [f"        case {i}: pat{tag}_{i}<R, RB>({call}); return true;" for i in range(len(keys))]
        out += ["        default: return false;", "    }", "}", ""]
        # host-side table of the same sequences, for the library's program checks
        offs, flat = [0], []
        for key in keys:
            flat += [list(op) for op in key]
            offs.append(len(flat))
        out.append(f"constexpr int kPatCount{tag} = {len(keys)};")
        out.append(f"constexpr int kPatOff{tag}[] = {{{', '.join(map(str, offs))}}};")
        rows = ", ".join("{" + ", ".join(map(str, op)) + "}" for op in flat) or "{0, 0, 0, 0, 0}"
        out.append(f"constexpr signed char kPatOps{tag}[][5] = {{{rows}}};")
        out.append("")
    out += ["}  // namespace svh", ""]
    table = {f"{rb}:{int(dual)}": [list(map(list, k)) for k in keys] for (rb, dual), keys in sorted(keys_by_lib.items())}
    return "\n".join(out), json.dumps(table, separators=(",", ":"))


def _canonical_programs():
    """(program, dual) pairs whose stage patterns make up the library: the
    hardware-efficient ansatz (RY, RZ on every qubit, then a CZ ladder) at
    10-22 qubits and 1-32 layers in the tile shapes the batched and
    single-circuit adjoint kernels use."""
    from dataclasses import replace

    import torch

    from .adjoint import ParamProgram
    from .batch import batch_config
    from .workloads import hea_ops, tfim_terms
    from .autotune import best_adjoint_config
    out = []
    for rb in LIB_RB:
        for n in range(10, 23):
            for L in range(1, 33):
                for remap in (True, False):
                    cfg, adj_cfg = batch_config(n, torch.complex64, rb, remap)
                    pp = ParamProgram(None, tfim_terms(n), 2 * n * L, cfg=cfg, adj_cfg=adj_cfg,
                                      dtype=torch.complex64, ops=hea_ops(n, L), n=n)
                    out += ([(pp.fwd, False)] if remap else []) + [(pp.adj, True)]
    for n in (16, 20, 22):
        cfg = replace(best_adjoint_config(n, torch.complex64), specialize=False, lds_split=0, generic=True)
        pp = ParamProgram(None, tfim_terms(n), 32 * n, cfg=cfg, dtype=torch.complex64, ops=hea_ops(n, 16), n=n)
        out += [(pp.fwd, False), (pp.adj, True)]
    return out


def build_library():
    from .specialize import _decode
    keys = {(rb, d): [] for rb in DISPATCH_RB for d in (False, True)}
    seen = {k: set() for k in keys}
    for prog, dual in _canonical_programs():
        _, passes, stages, ops = _decode(prog.blob)
        for p in passes:
            rb = int(p['rb'])
            if int(p['kind']) != P.PASS_TILE or rb not in LIB_RB:
                continue
            for s in range(int(p['stage_begin']), int(p['stage_begin']) + int(p['n_stages'])):
                st = stages[s]
                key = stage_key(ops[int(st['op_begin']):int(st['op_end'])], dual)
                if key and key not in seen[(rb, dual)]:
                    seen[(rb, dual)].add(key)
                    keys[(rb, dual)].append(key)
    return keys


def main():
    global _table
    keys = build_library()
    header, table = generate(keys)
    with open(HEADER, "w") as f:
        f.write(header)
    with open(TABLE, "w") as f:
        f.write(table)
    _table = None
    for k, v in sorted(keys.items()):
        print(f"rb={k[0]} dual={int(k[1])}: {len(v)} patterns, {sum(len(x) for x in v)} ops")


if __name__ == "__main__":
    main()
