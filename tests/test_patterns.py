"""Generic-kernel programs on the CPU: the compiler's output for generic
passes (interpreted by tests/prog_interp.py) matches the oracle, and every
stage's pattern id names exactly the stage's op sequence in the committed
library; the generated header and the key table agree."""
import json
import os
import re

import numpy as np
import pytest
import torch

import prog_interp
from oracle import sv_oracle as O
from qaravan_amd import patterns as PT
from qaravan_amd import program as P
from qaravan_amd.adjoint import ParamProgram
from qaravan_amd.specialize import _decode
from qaravan_amd.workloads import hea_ops, hea_theta, tfim_terms


def _check_ids(prog, dual):
    _, passes, stages, ops = _decode(prog.blob)
    lib = PT.load_table()
    inv = {k: {i: key for key, i in v.items()} for k, v in lib.items()}
    used = 0
    for p in passes:
        if int(p['kind']) != P.PASS_TILE:
            continue
        assert int(p['flags']) & P.PASSF_GENERIC
        s0, ns = int(p['stage_begin']), int(p['n_stages'])
        for s in range(s0, s0 + ns):
            pid = int(stages[s]['pattern'])
            key = PT.stage_key(ops[int(stages[s]['op_begin']):int(stages[s]['op_end'])], dual)
            if pid >= 0:
                assert inv[(int(p['rb']), int(dual))][pid] == key
                used += 1
            else:
                assert key not in lib.get((int(p['rb']), int(dual)), {}) or not key
        if int(p['flags']) & P.PASSF_REMAP_STORE:
            assert int(stages[s0 + ns]['pattern']) == -1
    return used


def test_header_matches_table():
    raw = json.load(open(PT.TABLE))
    text = open(PT.HEADER).read()
    for name, keys in raw.items():
        rb, d = name.split(":")
        tag = f"{rb}{'d' if d == '1' else 'f'}"
        assert f"constexpr int kPatCount{tag} = {len(keys)};" in text
        nfun = len(re.findall(rf"void pat{tag}_\d+\(", text))
        assert nfun == len(keys)


@pytest.mark.parametrize("n,layers,tb,rb,remap", [(10, 3, 6, 3, False), (13, 5, 8, 3, True), (11, 3, 7, 4, False)])
def test_generic_adjoint_programs_interpret_to_oracle(n, layers, tb, rb, remap):
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=1e9, rb=rb, generic=True, io_lane_bits=4, remap=remap)
    pp = ParamProgram(None, tfim_terms(n), 2 * n * layers, cfg=cfg, dtype=torch.complex64, ops=hea_ops(n, layers),
                      n=n)
    used = _check_ids(pp.fwd, False) + _check_ids(pp.adj, True)
    assert (used > 0) == (rb == 3)
    theta = hea_theta(n, layers)
    sv, _ = prog_interp.run(pp.fwd.blob, O.init_state(n), theta)      # ends in the program's layout
    if pp.layout is not None:
        sv = prog_interp.permute_bits(sv, n, {pp.layout[q]: q for q in range(n)})
    spec = []
    for l in range(layers):
        for k, w in enumerate("yz"):
            spec += [(w, [q], (2 * l + k) * n + q, None) for q in range(n)]
        spec += [(None, [i, i + 1], None, np.diag([1, 1, 1, -1]).astype(complex)) for i in range(n - 1)]
    ref = O.param_run(spec, theta, n)
    assert np.linalg.norm(sv - ref) < 1e-12


def test_generic_config_limits():
    ops = hea_ops(12, 1)
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=5, rb=3, generic=True))             # tb < 6
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=6, rb=5, generic=True))             # rb 5
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=6, rb=3, generic=True, lds_split=1))
