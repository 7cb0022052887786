"""CPU interpreter of encoded gate programs (test helper).

Decodes the byte layout of include/qaravan_hip.h and applies every pass,
stage and op in order with the oracle's ``op_action``, translating register
bits -> tile-local bits -> address bits -> qaravan qubits exactly as the HIP
kernels do.  It checks the compiler (lowering, fusion, partition, stages and
encoding) on the CPU; the kernels' own data movement is checked on the GPU.
"""
import numpy as np

from oracle import sv_oracle as O
from qaravan_amd import program as P

_CX_TC = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], dtype=np.complex128)
_CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
_WORD = {1: 'x', 2: 'y', 3: 'z'}


def decode(blob):
    h = np.frombuffer(blob[:P.HDR_DT.itemsize].tobytes(), dtype=P.HDR_DT)[0]

    def arr(name, dt, count):
        off = int(h['off_' + name])
        return np.frombuffer(blob[off:off + dt.itemsize * count].tobytes(), dtype=dt)
    passes = arr('passes', P.PASS_DT, int(h['n_passes']))
    stages = arr('stages', P.STAGE_DT, int(h['n_stages']))
    ops = arr('ops', P.OP_DT, int(h['n_ops']))
    grad = arr('grad', np.dtype('<i4'), int(h['n_grad_slots']))
    pool = arr('pool', np.dtype('<f8'), int(h['n_pool']))
    return h, passes, stages, ops, grad, pool


def _op_matrix(o, theta):
    kind = int(o['kind'])
    m = o['m'].view(np.complex128)
    if kind == P.OP_MAT1:
        return m[:4].reshape(2, 2)
    if kind == P.OP_MAT2:
        return m[:16].reshape(4, 4)
    if kind == P.OP_CX:
        return _CX_TC
    if kind == P.OP_CZ:
        return _CZ
    s = -1.0 if int(o['flags']) & P.OPF_NEG else 1.0
    w = _WORD[int(o['pauli'])]
    if kind == P.OP_ROT1:
        return O.rotation(w, s * theta[int(o['param'])])
    if kind == P.OP_ROT2:
        return O.rotation(w + w, s * theta[int(o['param'])])
    raise ValueError(kind)


def run(blob, sv, theta=None, lam=None, grad=None):
    """Apply the program to flat state ``sv`` (and ``lam`` in adjoint mode);
    adds 2·Im⟨λ|P|ψ⟩ into ``grad`` for flagged ops.  Returns (sv, lam)."""
    h, passes, stages, ops, gslots, pool = decode(blob)
    n = int(h['n_qubits'])
    sv = np.asarray(sv, dtype=np.complex128).copy()
    lam = None if lam is None else np.asarray(lam, dtype=np.complex128).copy()
    for p in passes:
        kind = int(p['kind'])
        if kind == P.PASS_NORMALIZE:
            sv = sv / np.linalg.norm(sv)
        elif kind == P.PASS_DENSE:
            k = int(p['L'])
            m = pool[int(p['pool_off']):int(p['pool_off']) + 2 * 4 ** k].view(np.complex128)
            qubits = [n - 1 - int(b) for b in p['bits'][:k]]
            sv = O.op_action(m.reshape(2 ** k, 2 ** k), qubits, sv)
        else:
            L = int(p['L'])
            tile_bits = [int(b) for b in p['bits'][:L]]
            if int(p['flags']) & P.PASSF_GEN_ZERO:
                # the pass starts from |0…0⟩ whatever the buffer holds
                sv = np.zeros_like(sv)
                sv[0] = 1.0
            for s in range(int(p['stage_begin']), int(p['stage_begin']) + int(p['n_stages'])):
                st = stages[s]
                rb = [int(x) for x in st['rbits'][:int(p['rb'])]]
                for oi in range(int(st['op_begin']), int(st['op_end'])):
                    o = ops[oi]
                    kind_o = int(o['kind'])
                    two = kind_o in (P.OP_MAT2, P.OP_CX, P.OP_CZ, P.OP_ROT2)
                    regs = [int(o['a'])] + ([int(o['b'])] if two else [])
                    qubits = [n - 1 - tile_bits[rb[r]] for r in regs]
                    if lam is not None and int(o['flags']) & P.OPF_GRAD:
                        w = _WORD[int(o['pauli'])]
                        pw = O.pauli_string_matrix(w if kind_o == P.OP_ROT1 else w + w)
                        slot = int(p['grad_begin']) + int(o['grad_slot'])
                        grad[gslots[slot]] += 2.0 * np.imag(np.vdot(lam, O.op_action(pw, qubits, sv)))
                    m = _op_matrix(o, theta)
                    sv = O.op_action(m, qubits, sv)
                    if lam is not None:
                        lam = O.op_action(m, qubits, lam)
            if int(p['flags']) & P.PASSF_REMAP_STORE:
                # the store record's tables give each tile bit its new address bit
                last = stages[int(p['stage_begin']) + int(p['n_stages']) - 1]
                rec = stages[int(p['stage_begin']) + int(p['n_stages'])]
                nrb, tb = int(p['rb']), int(p['tb'])
                rbits = [int(x) for x in last['rbits'][:nrb]]
                assert rbits == [int(x) for x in rec['rbits'][:nrb]]
                store = [None] * L
                for k, x in enumerate(rbits):
                    go = int(rec['goff'][1 << k])
                    assert go > 0 and go & (go - 1) == 0
                    store[x] = go.bit_length() - 1
                thr = [x for x in range(L) if x not in rbits]
                for i, x in enumerate(thr):
                    store[x] = int(rec['tgbit'][i])
                assert sorted(store) == sorted(tile_bits)
                move = {tile_bits[x]: store[x] for x in range(L)}
                sv = permute_bits(sv, n, move)
                if lam is not None:
                    lam = permute_bits(lam, n, move)
    if int(h['off_cx']):
        # CNOTs (control, target) on physical bits, in gate order, before the permutation
        off = int(h['off_cx'])
        cnt = int(np.frombuffer(blob[off:off + 4].tobytes(), dtype='<i4')[0])
        pairs = np.frombuffer(blob[off + 4:off + 4 + 8 * cnt].tobytes(), dtype='<i4').reshape(cnt, 2)
        idx = np.arange(2 ** n, dtype=np.int64)
        for c, t in pairs:
            sv = np.asarray(sv)[idx ^ (((idx >> int(c)) & 1) << int(t))]
    if int(h['off_perm']):
        off = int(h['off_perm'])
        perm = np.frombuffer(blob[off:off + 4 * n].tobytes(), dtype='<i4')
        sv = permute_bits(sv, n, {int(perm[q]): q for q in range(n)})
    return sv, lam


def permute_bits(sv, n, move):
    """State with address bit b moved to bit move[b] (bits not in ``move``
    stay); axis i of the rank-n view is address bit n-1-i."""
    src_of = {d: b for b, d in move.items()}
    axes = [n - 1 - src_of.get(n - 1 - k, n - 1 - k) for k in range(n)]
    return np.transpose(np.asarray(sv).reshape((2,) * n), axes).reshape(-1).copy()
