"""Host checks of the round-synchronous batch with 3 register bits per thread
and LDS exchanges in halves, the configuration of the 512-circuit benchmark
batch that the library keeps switched off (BatchedAdjoint raises for rb != 4).

Every launch-geometry fact the device code relies on is checked on the CPU for
every distinct pass of the batch's programs (forward and adjoint), against the
same sizing functions the planner uses (qaravan_amd.batch):

* every pass is one the compiled kernels take (specialize.eligible);
* every split exchange keeps the split bit in registers on both sides and,
  executed for every thread, delivers each tile element to the register the
  next layout expects through slots inside the half-size LDS buffer;
* the LDS the generated kernel text addresses (ψ buffer, λ buffer at its
  offset, one row of gradient slots per wave) fits the launch's dynamic LDS;
* the gradient-partial rows a pass writes (workgroup × n_grad) fit the
  circuit's partial buffer, and the reduction reads exactly those rows;
* every stage's address tables cover the pass's tile once, inside the
  circuit's 2^n amplitudes, and its LDS indices are a permutation of the
  2^L tile slots.
"""
import hashlib
import re

import numpy as np
import pytest
import torch

import prog_interp
from qaravan_amd import program as P
from qaravan_amd import specialize as SP
from qaravan_amd.batch import batch_programs, default_tb, gpart_rows, pass_grad_ops, spb_lds_bytes
from qaravan_amd.workloads import batch_specs, hea_ops, tfim_terms
from test_program import _run_split_exchange

C64 = torch.complex64


@pytest.fixture(scope="module")
def programs():
    ns, ls, thetas, _ = batch_specs(512, 20260004)
    keys = list(zip(ns.tolist(), ls.tolist()))
    first = {}
    for c, k in enumerate(keys):
        first.setdefault(k, c)
    idx = sorted(first.values())
    sel_ns = [int(ns[c]) for c in idx]
    pps = batch_programs(sel_ns, None, [hea_ops(int(ns[c]), int(ls[c])) for c in idx],
                         [tfim_terms(int(ns[c])) for c in idx], [len(thetas[c]) for c in idx], C64,
                         lambda n: default_tb(n, C64, 3), 1e9, None, 1, 3, False)
    return list(zip(sel_ns, pps))


def _stage_check(n, p, st, tb, rb):
    L = int(p['L'])
    bits = [int(b) for b in p['bits'][:L]]
    assert max(bits) < n
    tile = np.zeros(1, dtype=np.int64)
    for b in bits:
        tile = np.concatenate([tile, tile | (1 << b)])
    t = np.arange(1 << tb, dtype=np.int64)
    tg = np.zeros_like(t)
    for i in range(tb):
        assert int(st['tgbit'][i]) < n
        tg |= ((t >> i) & 1) << int(st['tgbit'][i])
    tl = t.copy()
    for r in st['rsorted'][:rb]:
        r = int(r)
        tl = ((tl >> r) << (r + 1)) | (tl & ((1 << r) - 1))
    swz = tl ^ ((tl >> 5) & 31)
    g = np.concatenate([tg | int(st['goff'][j]) for j in range(1 << rb)])
    l = np.concatenate([swz ^ int(st['lsw'][j]) for j in range(1 << rb)])
    assert np.array_equal(np.sort(g), np.sort(tile))
    assert np.array_equal(np.sort(l), np.arange(1 << L))


def test_rb3_split_batch_geometry(programs):
    es = 8
    seen_stage, seen_ex, seen_body = set(), set(), set()
    n_pass = n_ex = 0
    for n, pp in programs:
        rows = gpart_rows(pp, n)
        for dual, prog in ((False, pp.fwd), (True, pp.adj)):
            h, passes, stages, ops, _, _ = prog_interp.decode(prog.blob)
            for pi, p in enumerate(passes):
                if int(p['kind']) != P.PASS_TILE:
                    continue
                n_pass += 1
                tb, rb, L = int(p['tb']), int(p['rb']), int(p['L'])
                assert rb == 3 and tb == default_tb(n, C64, 3) and L == tb + rb
                assert SP.eligible(p, 0, dual)
                k = SP.lds_split(p, dual)
                ng = int(p['n_grad'])
                assert ng == (pass_grad_ops(prog.passes[pi]) if dual else 0)
                # rows of gradient partials: one per workgroup, ng slots each
                nwg = 2 ** (n - L)
                assert nwg * max(ng, 1) <= rows
                # LDS the generated text addresses vs the launch's dynamic LDS
                body = SP.batch_body(prog.blob, pi, 0, dual)
                key = hashlib.sha1(body.encode()).hexdigest()
                if key not in seen_body:
                    seen_body.add(key)
                    need = es * 2 ** (L - k)
                    if dual:
                        m2 = re.search(r"lds2 = lds \+ (\d+);", body)
                        mg = re.search(r"gacc = reinterpret_cast<double \*>\(lds2 \+ (\d+)\);", body)
                        assert m2 and mg and int(m2.group(1)) == 2 ** (L - k) and int(mg.group(1)) == 2 ** (L - k)
                        nw = max(1, 2 ** tb // 64)
                        slots = [int(x) for x in re.findall(r"gacc, (\d+)\);", body)]
                        assert all(x == ng for x in slots)
                        need = 2 * es * 2 ** (L - k) + 8 * nw * ng
                    assert need <= spb_lds_bytes(tb, rb, k, dual, es, ng) <= SP.MAX_LDS
                # gradient slots of this pass index rows [0, ng) of gacc
                s0, ns_ = int(p['stage_begin']), int(p['n_stages'])
                if dual:
                    for s in range(s0, s0 + ns_):
                        for o in ops[int(stages[s]['op_begin']):int(stages[s]['op_end'])]:
                            if int(o['flags']) & P.OPF_GRAD:
                                assert 0 <= int(o['grad_slot']) < ng
                # stage address tables
                for s in range(s0, s0 + ns_ + (1 if int(p['flags']) & P.PASSF_REMAP_STORE else 0)):
                    skey = (n, tuple(int(b) for b in p['bits'][:L]), stages[s].tobytes())
                    if skey not in seen_stage:
                        seen_stage.add(skey)
                        _stage_check(n, p, stages[s], tb, rb)
                # split exchanges
                for s in range(1, ns_):
                    Pst, Sst = stages[s0 + s - 1], stages[s0 + s]
                    hs = SP.split_bits(Pst['rbits'], Sst['rbits'], rb, k)
                    assert len(hs) == k
                    ekey = (tb, L, Pst['rbits'].tobytes(), Sst['rbits'].tobytes(), Pst['rsorted'].tobytes(),
                            Sst['rsorted'].tobytes(), s == 1)
                    if k and ekey not in seen_ex:
                        seen_ex.add(ekey)
                        lines = SP._split_exchange("float", rb, Pst['rbits'], Sst['rbits'], s == 1, s - 1, s, k)
                        _run_split_exchange(lines, tb, rb, Pst, Sst, L, hs)
                        n_ex += 1
    assert n_pass > 1000 and n_ex > 0
