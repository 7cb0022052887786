"""The circuit compiler checked on the CPU: programs interpreted op by op
(tests/prog_interp.py) must reproduce the oracle's state."""
import numpy as np
import pytest

import prog_interp
from oracle import sv_oracle as O
from qaravan_amd import CNOT, CZ, Circuit, Gate, H, Reset, X, random_unitary
from qaravan_amd import program as P
from qaravan_amd.workloads import brickwork_circuit, hea_circuit, hea_theta, tfim_terms


def random_circuit(n, n_gates, rng, kinds=("u1", "u2", "cx", "cz"), k3=False):
    gates = []
    for _ in range(n_gates):
        kind = kinds[int(rng.integers(len(kinds)))]
        if kind == "u1":
            gates.append(Gate("U", [int(rng.integers(n))], random_unitary(2, rng)))
        elif kind == "u2":
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(Gate("U", qs, random_unitary(4, rng)))
        elif kind == "cx":
            i = int(rng.integers(n - 1))
            gates.append(CNOT([i, i + 1] if rng.random() < 0.5 else [i + 1, i], n))
        elif kind == "cz":
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(CZ(qs))
    if k3:
        qs = [int(x) for x in rng.choice(n, size=3, replace=False)]
        gates.insert(len(gates) // 2, Gate("U3", qs, random_unitary(8, rng)))
    return Circuit(gates, n=n)


def compile_circ(circ, cfg):
    return P.compile_ops(P.lower_circuit(circ.build(), circ.num_sites), circ.num_sites, cfg)


def test_lowering_recognises_cnot_orientations_and_cz():
    n = 4
    ops = P.lower_circuit(Circuit([CNOT([1, 2], n), CNOT([3, 2], n), CZ([0, 3])], n=n).build(), n)
    kinds = [(o.kind, o.bits) for o in ops]
    # CNOT([1,2]): target qubit 1 (bit 2), control qubit 2 (bit 1)
    assert ('cx', (2, 1)) in kinds
    # CNOT([3,2]) is stored as indices [2,3] with the swapped matrix: target qubit 3 (bit 0)
    assert ('cx', (0, 1)) in kinds
    assert ('cz', (3, 0)) in kinds


@pytest.mark.parametrize("tb,low,cost", [(4, 6, 96.0), (4, 4, 24.0), (5, 6, 1e9), (6, 2, 40.0),
                                          (8, 6, 96.0)])
@pytest.mark.parametrize("seed", range(3))
def test_program_matches_oracle(tb, low, cost, seed):
    rng = np.random.default_rng(100 + seed)
    n = max(tb + 4, 9) + seed
    circ = random_circuit(n, 60, rng, k3=(seed == 1))
    prog = compile_circ(circ, P.TileConfig(tb=tb, low_bits=low, max_cost=cost))
    got, _ = prog_interp.run(prog.blob, O.init_state(n))
    ref = O.run(circ.gate_list, n)
    assert np.linalg.norm(got - ref) < 1e-12


def test_small_register_uses_dense_passes():
    circ = Circuit([H(0), CNOT([1, 0], 3), CNOT([2, 1], 3)], n=3)
    prog = compile_circ(circ, P.TileConfig())
    assert all(p.kind == P.PASS_DENSE for p in prog.passes)
    got, _ = prog_interp.run(prog.blob, O.init_state(3))
    assert np.allclose(got, O.run(circ.gate_list, 3))


def test_reset_lowering_normalises():
    gates = [H(0), CNOT([1, 0]), Reset(1, "1"), H(1), CNOT([2, 1])]
    circ = Circuit(gates, n=9)
    prog = compile_circ(circ, P.TileConfig(tb=4))
    assert any(p.kind == P.PASS_NORMALIZE for p in prog.passes)
    got, _ = prog_interp.run(prog.blob, O.init_state(9))
    assert np.allclose(got, O.run(gates, 9), atol=1e-12)


def test_brickwork_12q_program():
    circ = brickwork_circuit(12, 20)
    for cfg in (P.TileConfig(tb=8, low_bits=6, max_cost=96), P.TileConfig(tb=6, low_bits=4, max_cost=48)):
        prog = compile_circ(circ, cfg)
        got, _ = prog_interp.run(prog.blob, O.init_state(12))
        assert np.linalg.norm(got - O.run(circ.gate_list, 12)) < 1e-12


def test_every_stage_op_fits_its_register_bits():
    circ = brickwork_circuit(20, 4)
    prog = compile_circ(circ, P.TileConfig(tb=8, low_bits=6, max_cost=96))
    h, passes, stages, ops, _, _ = prog_interp.decode(prog.blob)
    for p in passes:
        if p['kind'] != P.PASS_TILE:
            continue
        bits = list(p['bits'][:p['L']])
        assert bits == sorted(bits) and len(set(bits)) == p['L']
        assert set(range(6)) <= set(bits)
        first = stages[p['stage_begin']]
        last = stages[p['stage_begin'] + p['n_stages'] - 1]
        # lanes stay on the six lowest tile bits for the HBM load and store
        assert not set(first['rbits'][:4]) & set(range(6))
        assert not set(last['rbits'][:4]) & set(range(6))
        for s in range(p['stage_begin'], p['stage_begin'] + p['n_stages']):
            st = stages[s]
            assert len(set(st['rbits'][:4])) == 4
            for o in ops[st['op_begin']:st['op_end']]:
                assert 0 <= o['a'] < 4


def test_gate_count_and_fusion_c1():
    circ = brickwork_circuit(28, 40)
    ops = P.fuse_1q(P.lower_circuit(circ.build(), 28))
    n1 = sum(o.kind == 'mat1' for o in ops)
    ncx = sum(o.kind == 'cx' for o in ops)
    assert ncx == 540
    assert n1 <= 28 * 40          # RX·RY·RZ per qubit per layer fused into one 2x2


@pytest.mark.parametrize("tb,low,cost,io", [(3, 4, 1e9, 4), (4, 4, 48.0, 6), (6, 5, 160.0, 4), (8, 4, 1e9, 4)])
@pytest.mark.parametrize("seed", range(3))
def test_five_register_bit_programs_match_oracle(tb, low, cost, io, seed):
    rng = np.random.default_rng(500 + seed)
    n = max(tb + 5, 10) + seed
    circ = random_circuit(n, 70, rng, k3=(seed == 2))
    cfg = P.TileConfig(tb=tb, low_bits=low, max_cost=cost, specialize=True, io_lane_bits=io, rb=5)
    prog = compile_circ(circ, cfg)
    h, passes, stages, ops, _, _ = prog_interp.decode(prog.blob)
    tile = [p for p in passes if p['kind'] == P.PASS_TILE]
    assert tile and all(int(p['rb']) == 5 and int(p['L']) == int(p['tb']) + 5 for p in tile)
    for p in tile:
        for si in range(int(p['stage_begin']), int(p['stage_begin']) + int(p['n_stages'])):
            st = stages[si]
            assert len(set(int(x) for x in st['rbits'][:5])) == 5
            for o in ops[int(st['op_begin']):int(st['op_end'])]:
                assert 0 <= int(o['a']) < 5
    got, _ = prog_interp.run(prog.blob, O.init_state(n))
    assert np.linalg.norm(got - O.run(circ.gate_list, n)) < 1e-12


def test_register_bit_options_are_checked():
    circ = brickwork_circuit(12, 2)
    ops = P.lower_circuit(circ.build(), 12)
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=6, rb=5))                      # needs specialize=True
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=6, rb=5, specialize=True), grad=True)
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=6, rb=2, specialize=True))
    with pytest.raises(ValueError):
        P.compile_ops(ops, 12, P.TileConfig(tb=6, rb=3))                      # compiled kernels only


def _ops_unitary(ops, n):
    """Full 2^n x 2^n matrix of an op list, built column by column from the
    oracle's gate application on basis states (bits -> qubits n-1-b)."""
    cols = []
    for k in range(2 ** n):
        sv = np.zeros(2 ** n, dtype=np.complex128)
        sv[k] = 1.0
        for op in ops:
            if op.kind in ('mat1', 'mat2'):
                m = op.mat
            elif op.kind == 'cx':
                m = P._CX_TC
            else:
                m = P._CZ
            sv = O.op_action(np.asarray(m, dtype=np.complex128), [n - 1 - b for b in op.bits], sv)
        cols.append(sv)
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("seed", range(6))
def test_fuse_2q_preserves_the_unitary(seed):
    rng = np.random.default_rng(300 + seed)
    n = 5
    circ = random_circuit(n, 40, rng, kinds=("u1", "u1", "u1", "u2", "cx", "cz"))
    ops = P.fuse_1q(P.lower_circuit(circ.build(), n))
    fused = P.fuse_2q(ops)
    assert len(fused) < len(ops)
    assert np.linalg.norm(_ops_unitary(fused, n) - _ops_unitary(ops, n)) < 1e-12
    ref = np.stack([O.run(circ.gate_list, n, init=np.eye(2 ** n, dtype=np.complex128)[k])
                    for k in range(2 ** n)], axis=1)
    assert np.linalg.norm(_ops_unitary(ops, n) - ref) < 1e-12


def test_fuse_2q_halves_c1_alu_cost():
    """Brickwork: each even-layer CNOT takes the 1-qubit gates before and after
    it on both of its qubits (4 x 8 FMA -> one 16-FMA 2-qubit product)."""
    circ = brickwork_circuit(28, 40)
    ops = P.fuse_1q(P.lower_circuit(circ.build(), 28))
    fused = P.fuse_2q(ops)
    cost = lambda ol: sum(P.OP_COST[o.kind] for o in ol)
    kinds = [o.kind for o in fused]
    assert kinds.count('mat2') == 20 * 14 and kinds.count('cx') == 20 * 13
    assert kinds.count('mat1') < 60
    assert cost(fused) < 0.56 * cost(ops)


def test_fuse_2q_program_matches_oracle():
    circ = brickwork_circuit(12, 20)
    for f2 in (False, True):
        prog = compile_circ(circ, P.TileConfig(tb=6, low_bits=4, max_cost=48, fuse2q=f2))
        got, _ = prog_interp.run(prog.blob, O.init_state(12))
        assert np.linalg.norm(got - O.run(circ.gate_list, 12)) < 1e-12


def _param_ops(circ, n):
    return P.lower_circuit(circ.build(), n, param_mode=True)


@pytest.mark.parametrize("tb,layers,cost,rb,remap", [(4, 3, 40, 4, False), (6, 3, 40, 4, False),
                                                     (6, 6, 1e9, 4, False), (4, 5, 1e9, 4, False),
                                                     (5, 4, 1e9, 3, False), (7, 3, 40, 3, False),
                                                     (5, 6, 1e9, 3, True), (4, 8, 40, 4, True)])
def test_param_forward_and_adjoint_programs(tb, layers, cost, rb, remap):
    """Forward program with device-θ rotations, and the adjoint program's
    gradient, interpreted on the CPU, against the oracle's adjoint (RZ and CZ
    are diagonal, so the scheduler may reorder them).  With remap, both
    programs move qubits between address bits; the forward permutes back at
    the end, the adjoint sweep leaves ψ and λ permuted."""
    from qaravan_amd.adjoint import ParamProgram, term_table
    n = 10
    theta = hea_theta(n, layers)
    circ = hea_circuit(n, layers, theta)
    terms = tfim_terms(n)
    pp = ParamProgram(circ, terms, len(theta), cfg=P.TileConfig(tb=tb, low_bits=4 if remap else 6, max_cost=cost,
                                                                rb=rb, specialize=rb != 4 or remap, remap=remap,
                                                                io_lane_bits=4 if remap else 6))
    if remap:
        assert any(p.flags & P.PASSF_REMAP_STORE for p in pp.adj.passes)
        assert pp.adj.final_pos is None
        # the forward leaves ψ in its own layout, with no permutation back
        assert pp.layout is not None and pp.fwd.final_pos is None
    else:
        assert pp.layout is None
    assert all(p.rb == rb for p in pp.adj.passes if p.kind == P.PASS_TILE)
    psi, _ = prog_interp.run(pp.fwd.blob, O.init_state(n), theta=theta)
    ref = O.run(circ.gate_list, n)
    to_nat = {} if pp.layout is None else {pp.layout[q]: q for q in range(n)}
    assert np.linalg.norm(prog_interp.permute_bits(psi, n, to_nat) - ref) < 1e-12
    # λ = Hψ in the same layout, as the device forms it from the relabelled term table
    lam = prog_interp.permute_bits(O.apply_pauli_sum(ref, terms), n, {q: p for p, q in to_nat.items()})
    assert np.allclose(lam, _apply_terms_table(term_table(n, terms, pp.layout), psi), atol=1e-12)
    grad = np.zeros(len(theta))
    prog_interp.run(pp.adj.blob, psi, theta=theta, lam=lam, grad=grad)
    from test_oracle import hea_spec
    _, gref = O.adjoint_gradient(hea_spec(n, layers), theta, n, terms)
    assert np.allclose(grad, gref, atol=1e-10)


def _apply_terms_table(tab, psi):
    """λ_i = Σ_t c_t i^{ny} (-1)^{pc((i^x)&z)} ψ[i^x] from a device term table
    (the kernel's formula, restated), for checking relabelled tables."""
    idx = np.arange(len(psi), dtype=np.int64)
    lam = np.zeros_like(psi)
    for t in tab:
        j = idx ^ int(t['x'])
        sg = 1 - 2 * (np.array([bin(v).count("1") for v in (j & int(t['z']))]) & 1)
        lam += (t['cre'] + 1j * t['cim']) * sg * psi[j]
    return lam


def _diag_rich_circuit(n, n_gates, rng):
    """Random gates where most are diagonal (CZ, RZ, diagonal 1q/2q matrices)."""
    from qaravan_amd import RZ
    gates = []
    for _ in range(n_gates):
        r = rng.random()
        q = int(rng.integers(n))
        if r < 0.25:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(CZ(qs))
        elif r < 0.45:
            gates.append(RZ([q], float(rng.uniform(0, 6))))
        elif r < 0.6:
            gates.append(Gate("D", [q], np.diag(np.exp(1j * rng.uniform(0, 6, size=2)))))
        elif r < 0.72:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            gates.append(Gate("D2", qs, np.diag(np.exp(1j * rng.uniform(0, 6, size=4)))))
        elif r < 0.86:
            gates.append(Gate("U", [q], random_unitary(2, rng)))
        else:
            i = int(rng.integers(n - 1))
            gates.append(CNOT([i, i + 1], n))
    return Circuit(gates, n=n)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tb,rb", [(4, 4), (6, 4), (5, 5), (7, 3)])
def test_diagonal_reordering_matches_oracle(seed, tb, rb):
    rng = np.random.default_rng(700 + seed)
    n = 11
    circ = _diag_rich_circuit(n, 150, rng)
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=1e9, specialize=rb != 4, rb=rb)
    prog = compile_circ(circ, cfg)
    got, _ = prog_interp.run(prog.blob, O.init_state(n))
    assert np.linalg.norm(got - O.run(circ.gate_list, n)) < 1e-12


def test_diagonal_ops_cut_hea_passes():
    """The 20-qubit HEA of the gradient workload: RZ layers and the CZ ladder
    commute, so far fewer passes are needed than with strict gate order."""
    from qaravan_amd.adjoint import ParamProgram
    n, layers = 20, 16
    theta = hea_theta(n, layers)
    pp = ParamProgram(hea_circuit(n, layers, theta), tfim_terms(n), len(theta),
                      cfg=P.TileConfig(tb=8, low_bits=4, max_cost=1e9))
    assert pp.fwd.n_tile_passes() <= 24 and pp.adj.n_tile_passes() <= 24


def test_stage_tables_cover_each_tile_once():
    """Emulates the kernel's address arithmetic for every stage: the global
    offsets (thread part | register part) hit every amplitude of the tile once,
    and the swizzled LDS indices are a permutation of [0, 2^L)."""
    circ = brickwork_circuit(16, 3)
    for cfg in (P.TileConfig(tb=4, low_bits=4, max_cost=40), P.TileConfig(tb=8, low_bits=6, max_cost=96),
                P.TileConfig(tb=6, low_bits=4, max_cost=1e9, io_lane_bits=4, rb=5, specialize=True),
                P.TileConfig(tb=7, low_bits=4, max_cost=1e9, io_lane_bits=3, rb=3, specialize=True)):
        prog = compile_circ(circ, cfg)
        h, passes, stages, ops, _, _ = prog_interp.decode(prog.blob)
        for p in passes:
            if p['kind'] != P.PASS_TILE:
                continue
            L, tb, rb = int(p['L']), int(p['tb']), int(p['rb'])
            tile = set()
            for i in range(1 << L):
                a = 0
                for k in range(L):
                    a |= ((i >> k) & 1) << int(p['bits'][k])
                tile.add(a)
            for s in range(p['stage_begin'], p['stage_begin'] + p['n_stages']):
                st = stages[s]
                seen_g, seen_l = set(), set()
                for t in range(1 << tb):
                    tg = 0
                    for i in range(tb):
                        tg |= ((t >> i) & 1) << int(st['tgbit'][i])
                    tl = t
                    for r in st['rsorted'][:rb]:
                        r = int(r)
                        tl = ((tl >> r) << (r + 1)) | (tl & ((1 << r) - 1))
                    for j in range(1 << rb):
                        seen_g.add(tg | int(st['goff'][j]))
                        seen_l.add(P.lds_swz(tl) ^ int(st['lsw'][j]))
                assert seen_g == tile
                assert seen_l == set(range(1 << L))


def _thr_local(t, rsorted):
    for r in rsorted:
        r = int(r)
        t = ((t >> r) << (r + 1)) | (t & ((1 << r) - 1))
    return t


def _run_split_exchange(lines, tb, rb, Pst, Sst, L, hs):
    """Executes one generated split exchange on the CPU for every thread:
    thread t starts with register j = tile element thr_local_P(t) ^ loc_P(j)
    and must end with element thr_local_S(t) ^ loc_S(j); every slot lies in
    the 2^(L-k) LDS buffer.  (The thread part of a slot follows the same
    formula as the generated code: swizzle of the thread index with the split
    bits removed.)"""
    import re
    from qaravan_amd import specialize as SP

    def slot_thread(tl):
        return P.lds_swz(SP.drop_bits(tl, hs))

    def loc(rbits, j):
        return sum(((j >> k) & 1) << int(rbits[k]) for k in range(rb))

    for h in hs:
        assert f">> {h + 1}) << {h})" in lines[1]
    nt = 1 << tb
    v = [[_thr_local(t, Pst['rsorted'][:rb]) ^ loc(Pst['rbits'], j) for j in range(1 << rb)] for t in range(nt)]
    w = [[None] * (1 << rb) for _ in range(nt)]
    lds = [None] * (1 << (L - len(hs)))
    for ln in lines:
        m = re.match(r"\s+lds\[tp \^ (\d+)\] = v\[(\d+)\];", ln)
        if m:
            c, j = int(m.group(1)), int(m.group(2))
            for t in range(nt):
                i = slot_thread(_thr_local(t, Pst['rsorted'][:rb])) ^ c
                assert 0 <= i < len(lds)
                lds[i] = v[t][j]
            continue
        m = re.match(r"\s+x\[(\d+)\] = lds\[tn \^ (\d+)\];", ln)
        if m:
            j, c = int(m.group(1)), int(m.group(2))
            for t in range(nt):
                i = slot_thread(_thr_local(t, Sst['rsorted'][:rb])) ^ c
                w[t][j] = lds[i]
            continue
    for t in range(nt):
        for j in range(1 << rb):
            assert w[t][j] == _thr_local(t, Sst['rsorted'][:rb]) ^ loc(Sst['rbits'], j), (t, j)


@pytest.mark.parametrize("n,tb,rb,k", [(16, 6, 5, 1), (14, 5, 4, 1), (15, 6, 3, 1), (16, 6, 5, 2), (15, 5, 4, 2)])
def test_split_lds_exchanges_move_every_element(n, tb, rb, k):
    """Every generated split LDS exchange (lds_split passes) delivers each
    tile element to the register the next layout expects, through slots
    inside the 1/2^k-size LDS buffer."""
    from qaravan_amd import specialize as SP
    circ = brickwork_circuit(n, 6)
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=1e9, io_lane_bits=4, rb=rb, specialize=True, lds_split=k)
    prog = compile_circ(circ, cfg)
    h, passes, stages, ops, _, _ = prog_interp.decode(prog.blob)
    checked = 0
    for p in passes:
        if p['kind'] != P.PASS_TILE or SP.lds_split(p, False) != k:
            continue
        s0, ns = int(p['stage_begin']), int(p['n_stages'])
        for s in range(1, ns):
            Pst, Sst = stages[s0 + s - 1], stages[s0 + s]
            hs = SP.split_bits(Pst['rbits'], Sst['rbits'], rb, k)
            assert len(hs) == k
            lines = SP._split_exchange("float", rb, Pst['rbits'], Sst['rbits'], s == 1, s - 1, s, k)
            assert sum("__syncthreads" in ln for ln in lines) == 2 * (1 << k) - (s == 1)
            _run_split_exchange(lines, int(p['tb']), rb, Pst, Sst, int(p['L']), hs)
            # the adjoint form moves λ (w, lds2, y) through the same slots at the same steps
            dual = SP._split_exchange("float", rb, Pst['rbits'], Sst['rbits'], s == 1, s - 1, s, k, dual=True)
            single = [ln for ln in dual if "lds2" not in ln and "y[" not in ln and " w[" not in ln]
            assert single == lines
            mirrored = [ln.replace("lds2", "lds").replace("= w[", "= v[").replace("y[", "x[")
                        for ln in dual if "lds2" in ln]
            assert mirrored == [ln for ln in lines if "lds[" in ln]
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("k", [1, 2])
def test_lds_split_flag_and_shared_bits(k):
    circ = brickwork_circuit(16, 8)
    cfg = P.TileConfig(tb=6, low_bits=4, max_cost=160, io_lane_bits=4, rb=5, specialize=True, lds_split=k)
    prog = compile_circ(circ, cfg)
    nflag = 0
    for p in prog.passes:
        if p.kind != P.PASS_TILE:
            continue
        shared = min((len(set(a.rbits) & set(b.rbits)) for a, b in zip(p.stages, p.stages[1:])), default=5)
        want = min(k, shared)   # a pass that cannot keep k shared bits splits in fewer parts
        got = 1 if p.flags & P.PASSF_HALF_LDS else 2 if p.flags & P.PASSF_QUARTER_LDS else 0
        assert got == want
        nflag += got == k
        # first and last layouts keep the 4 lane bits off the registers
        assert not set(p.stages[0].rbits) & set(range(4))
        assert not set(p.stages[-1].rbits) & set(range(4))
    assert nflag > 0
    with pytest.raises(ValueError):
        compile_circ(circ, P.TileConfig(tb=6, low_bits=4, rb=4, lds_split=k))
    with pytest.raises(ValueError):
        compile_circ(circ, P.TileConfig(tb=6, low_bits=4, rb=3, specialize=True, lds_split=3))


@pytest.mark.parametrize("n,layers,tb,cap,seed", [(12, 4, 5, 64, 0), (14, 8, 6, 1e9, 1), (16, 10, 6, 96, 2),
                                                  (16, 6, 7, 160, 3)])
def test_remap_programs_match_oracle(n, layers, tb, cap, seed):
    """Programs that move qubits between address bits at each store (and
    permute back at the end) give the oracle's state, for the brickwork and
    for a random circuit; stores stay coalesced and cover each tile once."""
    rng = np.random.default_rng(seed)
    cfg = P.TileConfig(tb=tb, low_bits=4, max_cost=cap, specialize=True, io_lane_bits=4, rb=5, lds_split=1,
                       remap=True)
    for circ in (brickwork_circuit(n, layers), _random_tile_circuit(n, 80, rng)):
        prog = compile_circ(circ, cfg)
        assert prog.final_pos is not None
        h, passes, stages, ops, _, _ = prog_interp.decode(prog.blob)
        nremap = 0
        for p in passes:
            if not int(p['flags']) & P.PASSF_REMAP_STORE:
                continue
            nremap += 1
            L, ptb, rb = int(p['L']), int(p['tb']), int(p['rb'])
            rec = stages[int(p['stage_begin']) + int(p['n_stages'])]
            assert [int(x) for x in rec['tgbit'][:4]] == [0, 1, 2, 3]      # 128-byte store runs
            tile = {sum(((i >> k) & 1) << int(p['bits'][k]) for k in range(L)) for i in range(1 << L)}
            seen = set()
            for t in range(1 << ptb):
                tg = sum(((t >> i) & 1) << int(rec['tgbit'][i]) for i in range(ptb))
                for j in range(1 << rb):
                    seen.add(tg | int(rec['goff'][j]))
            assert seen == tile
        assert nremap > 0
        sv0 = np.zeros(2 ** n, complex)
        sv0[0] = 1
        sv, _ = prog_interp.run(prog.blob, sv0)
        assert np.linalg.norm(sv - O.run(circ.gate_list, n)) < 1e-12


def _trailing_cx_circuits(n, seeds):
    """Brickwork and random circuits followed by a chain of CNOTs on far-apart
    qubits (non-commuting: each shares a qubit with the next), so that some
    remapped programs end in a pass of CNOTs only."""
    out = []
    for seed in seeds:
        rng = np.random.default_rng(seed)
        base = brickwork_circuit(n, 4) if seed % 2 == 0 else _random_tile_circuit(n, 60, rng)
        qs = [int(x) for x in rng.permutation(n)[:5]]
        cxs = [CNOT([qs[i], qs[i + 1]], n) for i in range(4)]
        out.append(Circuit(list(base.gate_list) + cxs, n=n))
    return out


def test_trailing_cnots_fold_into_final_permutation():
    """A remapped program whose last pass holds only CNOTs drops that pass
    and hands its CNOTs (physical bits, gate order) to the final
    permutation; the result is the oracle's state, and at least one of the
    circuits folds."""
    n = 12
    cfg = P.TileConfig(tb=5, low_bits=4, max_cost=1e9, specialize=True, io_lane_bits=4, rb=5, lds_split=1,
                       remap=True)
    folded = 0
    for circ in _trailing_cx_circuits(n, range(12)):
        prog = compile_circ(circ, cfg)
        h, passes, _, _, _, _ = prog_interp.decode(prog.blob)
        if prog.final_cx:
            folded += 1
            assert int(h['off_cx']) and int(h['off_perm'])
            assert 1 <= len(prog.final_cx) <= P.MAX_FINAL_CX
        else:
            assert int(h['off_cx']) == 0
        sv0 = np.zeros(2 ** n, complex)
        sv0[0] = 1
        sv, _ = prog_interp.run(prog.blob, sv0)
        assert np.linalg.norm(sv - O.run(circ.gate_list, n)) < 1e-12
    assert folded > 0


def test_final_cnots_need_a_permutation():
    with pytest.raises(ValueError):
        P.encode(10, [], final_pos=None, final_cx=[(0, 1)])


def _random_tile_circuit(n, ng, rng):
    gates = []
    for _ in range(ng):
        if rng.random() < 0.5:
            q = int(rng.integers(n))
            gates.append(Gate("U", [q], random_unitary(2, rng)))
        else:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            gates.append(CNOT([a, b], n) if rng.random() < 0.5 else Gate("U2", [a, b], random_unitary(4, rng)))
    return Circuit(gates, n=n)


def test_remap_falls_back_without_tile_ops():
    circ = Circuit([H(0), CNOT([1, 0], 10), Reset(3, "0"), CNOT([2, 1], 10)], n=10)
    cfg = P.TileConfig(tb=5, low_bits=4, max_cost=64, specialize=True, io_lane_bits=4, rb=5, remap=True)
    prog = compile_circ(circ, cfg)
    assert prog.final_pos is None
    assert not any(p.flags & P.PASSF_REMAP_STORE for p in prog.passes)


@pytest.mark.parametrize("remap2", [True, False])
def test_layout_handoff_between_programs(remap2):
    """A remapped program compiled with keep_layout leaves the state in its
    final layout (end_pos, no permutation back); a second program compiled
    from that layout (pos0) continues the circuit; together they give the
    oracle's state once the bits are moved back by the last layout."""
    n = 12
    c1 = brickwork_circuit(n, 3)
    rng = np.random.default_rng(3)
    c2 = _random_tile_circuit(n, 40, rng)
    cfg = P.TileConfig(tb=5, low_bits=4, max_cost=1e9, specialize=True, io_lane_bits=4, rb=5, lds_split=1,
                       remap=True)
    p1 = P.compile_ops(P.lower_circuit(c1.build(), n), n, cfg, keep_layout=True)
    assert p1.end_pos is not None and p1.final_pos is None
    cfg2 = cfg if remap2 else P.TileConfig(tb=5, low_bits=4, max_cost=1e9, rb=4, io_lane_bits=4)
    p2 = P.compile_ops(P.lower_circuit(c2.build(), n), n, cfg2, keep_layout=True, pos0=p1.end_pos)
    end = p2.end_pos or list(range(n))
    sv, _ = prog_interp.run(p1.blob, O.init_state(n))
    sv, _ = prog_interp.run(p2.blob, sv)
    sv = prog_interp.permute_bits(sv, n, {end[q]: q for q in range(n)})
    ref = O.run(list(c1.gate_list) + list(c2.gate_list), n)
    assert np.linalg.norm(sv - ref) < 1e-12
    with pytest.raises(ValueError):
        P.compile_ops(P.lower_circuit(c2.build(), n), n, cfg2, pos0=p1.end_pos)


def test_zero_start_variant_ignores_the_buffer():
    """The variant whose first pass generates |0…0⟩ gives the oracle's state
    whatever the state buffer held; the flag is refused anywhere but on the
    first pass of a forward program (host check, no GPU)."""
    import ctypes

    from qaravan_amd import _lib
    n = 12
    circ = brickwork_circuit(n, 3)
    cfg = P.TileConfig(tb=5, low_bits=4, max_cost=1e9, specialize=True, io_lane_bits=4, rb=5, lds_split=1,
                       remap=True)
    prog = P.compile_ops(P.lower_circuit(circ.build(), n), n, cfg)
    z = P.zero_start_variant(prog, 0)
    assert z is not None and z.final_pos == prog.final_pos
    junk = np.random.default_rng(0).standard_normal(2 ** n) + 0j
    sv, _ = prog_interp.run(z.blob, junk)
    assert np.linalg.norm(sv - O.run(circ.gate_list, n)) < 1e-12
    # the same flag on the second pass is refused by the library's program check
    h = np.frombuffer(z.blob[:P.HDR_DT.itemsize].tobytes(), dtype=P.HDR_DT)[0]
    bad = z.blob.copy()
    off = int(h['off_passes']) + P.PASS_DT.itemsize
    row = np.frombuffer(bad[off:off + P.PASS_DT.itemsize].tobytes(), dtype=P.PASS_DT)[0].copy()
    row['flags'] = int(row['flags']) | P.PASSF_GEN_ZERO
    bad[off:off + P.PASS_DT.itemsize] = np.frombuffer(row.tobytes(), dtype=np.uint8)
    lib = _lib.load()
    need = ctypes.c_size_t()
    assert lib.svh_program_scratch_bytes(z.blob.ctypes.data, 0, ctypes.byref(need)) == 0
    assert lib.svh_program_scratch_bytes(bad.ctypes.data, 0, ctypes.byref(need)) == -8
