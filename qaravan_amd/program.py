"""Circuit compiler: qaravan gate lists -> gate programs for the HIP kernels.

The reference applies every gate as its own tensor contraction, one Python call
per gate (src/qaravan/core/base_sim.py:40-47 -> tensorQ/statevector_sim.py:
124-140).  Here the whole layered circuit is lowered once into a program that
the library runs as a handful of fused passes:

1. Lowering.  Each gate becomes an elementary op on ADDRESS bits (qaravan qubit
   q is bit n-1-q, because the state is a C-order rank-n tensor, i.e.
   big-endian).  2-qubit matrices equal to a CNOT (either orientation) or to CZ
   become permutation / sign ops; other 1- and 2-qubit matrices stay dense;
   3..6-qubit gates become stand-alone dense passes; ``Reset`` becomes its
   projector followed by a renormalisation pass (base_sim.py:25-28).
2. Fusion.  Runs of dense 1-qubit ops on the same bit (with no other op on
   that bit in between) are multiplied into one 2x2; then 1-qubit ops are
   folded into an adjacent 2-qubit op where that saves ALU work (fuse_2q).
3. Pass partition.  Ops are packed greedily, in order, into tile passes: a pass
   owns L address bits (always the lowest ``low_bits`` ones, so every tile
   load is a run of contiguous amplitudes) and takes every op whose bits fit,
   stopping at an ALU-cost cap.  An op that does not fit blocks its bits for
   the rest of the pass, so the order of overlapping gates is preserved,
   except between two diagonal ops (CZ, Z-type rotations, diagonal matrices),
   which commute and may pass each other.
4. Stages.  Inside a pass, ops are grouped into stages of 4 register bits;
   each stage change is one LDS re-distribution in the kernel.

The byte layout matches include/qaravan_hip.h (svh_prog_header_t, svh_pass_t,
svh_stage_t, svh_op_t).
"""
from dataclasses import dataclass, field

import numpy as np

MAGIC = 0x53564850
ABI_VERSION = 1
PASS_TILE, PASS_DENSE, PASS_NORMALIZE = 1, 2, 3
PASSF_PREFETCH = 1
PASSF_HALF_LDS = 2
PASSF_QUARTER_LDS = 4
PASSF_REMAP_STORE = 8
PASSF_GENERIC = 16
PASSF_GEN_ZERO = 32
OP_MAT1, OP_MAT2, OP_CX, OP_CZ, OP_ROT1, OP_ROT2 = 1, 2, 3, 4, 5, 6
OPF_NEG, OPF_GRAD = 1, 2
RB = 4
MAX_TILE_BITS = 16

HDR_DT = np.dtype([('magic', '<u4'), ('version', '<i4'), ('n_qubits', '<i4'), ('n_passes', '<i4'),
                   ('n_stages', '<i4'), ('n_ops', '<i4'), ('n_grad_slots', '<i4'), ('n_pool', '<i4'),
                   ('off_passes', '<i8'), ('off_stages', '<i8'), ('off_ops', '<i8'),
                   ('off_grad', '<i8'), ('off_pool', '<i8'), ('n_params', '<i8'),
                   ('off_perm', '<i8'), ('off_cx', '<i8')])
MAX_FINAL_CX = 8
PASS_DT = np.dtype([('kind', '<i4'), ('L', '<i4'), ('tb', '<i4'), ('rb', '<i4'),
                    ('stage_begin', '<i4'), ('n_stages', '<i4'), ('grad_begin', '<i4'),
                    ('n_grad', '<i4'), ('pool_off', '<i8'), ('flags', '<i4'), ('pad', '<i4'),
                    ('bits', '<i4', (MAX_TILE_BITS,))])
STAGE_DT = np.dtype([('goff', '<i8', (32,)), ('lsw', '<i4', (32,)), ('rbits', '<i4', (8,)),
                     ('rsorted', '<i4', (8,)), ('tgbit', '<i4', (12,)), ('op_begin', '<i4'),
                     ('op_end', '<i4'), ('pattern', '<i4'), ('pad', '<i4')])
OP_DT = np.dtype([('kind', '<i4'), ('a', '<i4'), ('b', '<i4'), ('param', '<i4'), ('pauli', '<i4'),
                  ('flags', '<i4'), ('grad_slot', '<i4'), ('pad', '<i4'), ('m', '<f8', (32,)),
                  ('mf', '<f4', (32,))])
assert HDR_DT.itemsize == 96 and PASS_DT.itemsize == 112
assert STAGE_DT.itemsize == 512 and OP_DT.itemsize == 416


_ZERO_M = np.zeros(32)


def lds_swz(li):
    """The kernel's LDS swizzle (linear over GF(2)): low 5 bits ^= next 5."""
    return li ^ ((li >> 5) & 31)


_JBITS = {r: np.array([[(j >> k) & 1 for k in range(r)] for j in range(1 << r)], dtype=np.int64)
          for r in (3, 4, 5)}


def stage_tables(tile_bits, tb, rbits):
    """Address tables of one stage layout (see svh_stage_t): thread bits fill
    the non-register tile-local bits in ascending order; len(rbits) register
    bits give 2^len(rbits) registers."""
    r = len(rbits)
    L = tb + r
    rset = set(rbits)
    tgbit = [tile_bits[x] for x in range(L) if x not in rset]
    rb = np.asarray(rbits, dtype=np.int64)
    loc = _JBITS[r] @ (np.int64(1) << rb)
    goff = _JBITS[r] @ (np.int64(1) << np.asarray([tile_bits[x] for x in rbits], dtype=np.int64))
    return dict(rsorted=sorted(rbits), tgbit=tgbit, lsw=loc ^ ((loc >> 5) & 31), goff=goff, loc=loc)

PAULI_CODE = {'x': 1, 'y': 2, 'z': 3}

# per-amplitude VALU cost of an op inside a tile pass (used for the pass cap)
OP_COST = {'mat1': 8.0, 'rot1': 4.0, 'mat2': 16.0, 'cx': 0.5, 'cz': 0.25, 'rot2': 8.0}

_CX_TC = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]])  # target = row MSB
_CX_CT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])  # control = row MSB
_CZ = np.diag([1, 1, 1, -1])


@dataclass
class EOp:
    """Elementary op on address bits.

    kind: 'mat1' | 'mat2' | 'cx' | 'cz' | 'rot1' | 'rot2' | 'dense' | 'norm'
    bits: address bits; mat2/dense: bits[0] is the row MSB; cx: (target, control)
    """
    kind: str
    bits: tuple
    mat: np.ndarray = None
    param: int = -1
    pauli: int = 0
    flags: int = 0

    def dagger(self):
        if self.kind in ('mat1', 'mat2', 'dense'):
            return EOp(self.kind, self.bits, self.mat.conj().T.copy(), self.param, self.pauli, self.flags)
        if self.kind in ('rot1', 'rot2'):
            return EOp(self.kind, self.bits, None, self.param, self.pauli, self.flags ^ OPF_NEG)
        if self.kind in ('cx', 'cz'):
            return self
        raise ValueError(f"{self.kind} has no adjoint")


# ----------------------------------------------------------------- lowering
def _as_np(m):
    """A gate matrix as complex128 numpy.  A torch matrix that requires grad is
    refused while grad mode is on: a compiled program holds plain numbers, so
    the gradient would be dropped without a word.  Differentiable runs go
    through qaravan_amd.state_autograd (StatevectorSim(backend="torch")),
    which lowers detached values itself and supplies the backward pass."""
    try:
        import torch
        if torch.is_tensor(m):
            if m.requires_grad and torch.is_grad_enabled():
                raise NotImplementedError(
                    "gate matrix requires grad: compiling it into a program would drop the gradient; run the "
                    "circuit with StatevectorSim(backend='torch') (or under torch.no_grad() to use the values)")
            m = m.detach().resolve_conj().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(m, dtype=np.complex128)


def lower_matrix_gate(mat, qubits, n, name=None):
    """Ops for a gate given as a matrix over ``qubits`` (first = row MSB)."""
    mat = _as_np(mat)
    k = len(qubits)
    if mat.shape != (2 ** k, 2 ** k):
        raise ValueError(f"gate matrix shape {mat.shape} does not match {k} qubit(s)")
    if len(set(qubits)) != k:
        raise ValueError(f"repeated qubit in {qubits}")
    for q in qubits:
        if not 0 <= q < n:
            raise ValueError(f"qubit {q} outside a {n}-qubit register")
    bits = tuple(n - 1 - q for q in qubits)
    out = []
    if k == 1:
        out.append(EOp('mat1', bits, mat))
    elif k == 2:
        if np.array_equal(mat, _CX_TC):
            out.append(EOp('cx', (bits[0], bits[1])))
        elif np.array_equal(mat, _CX_CT):
            out.append(EOp('cx', (bits[1], bits[0])))
        elif np.array_equal(mat, _CZ):
            out.append(EOp('cz', bits))
        else:
            out.append(EOp('mat2', bits, mat))
    elif k <= 6:
        out.append(EOp('dense', bits, mat))
    else:
        raise NotImplementedError("gates on more than 6 qubits are not supported")
    if name == "Reset":
        pass  # normalisation appended by the caller when requested
    return out


def lower_circuit(circ, n=None, param_mode=False):
    """Lower a built qaravan ``Circuit`` (greedy layers, circuits.py:42-63) to
    EOps in layered order.  With ``param_mode``, rotations RX/RY/RZ/RXX/RYY/RZZ
    whose angle is given by ``gate.param_index`` become device-θ ops."""
    n = circ.num_sites if n is None else n
    if circ.local_dim != 2:
        raise NotImplementedError("the HIP statevector path supports qubits (local_dim = 2) only")
    layers = circ.layers if circ.layers is not None else circ.build().layers
    ops = []
    for layer in layers:
        for g in layer:
            ops.extend(lower_gate(g, n, param_mode))
    return ops


def lower_gate(g, n, param_mode=False):
    if param_mode and getattr(g, "param_index", None) is not None:
        word = getattr(g, "generator_word", None)
        bits = tuple(n - 1 - q for q in g.indices)
        if word in ('x', 'y', 'z'):
            return [EOp('rot1', bits, param=int(g.param_index), pauli=PAULI_CODE[word])]
        if word in ('xx', 'yy', 'zz'):
            return [EOp('rot2', bits, param=int(g.param_index), pauli=PAULI_CODE[word[0]])]
        raise NotImplementedError(f"{g.name} cannot take a device angle")
    ops = lower_matrix_gate(g.matrix, list(g.indices), n, g.name)
    if g.name == "Reset" and getattr(g, "normalize", True):
        ops.append(EOp('norm', ()))
    return ops


def fuse_1q(ops):
    """Multiply runs of dense 1-qubit ops per bit; flush a bit's pending matrix
    before any other op touches it."""
    pending = {}
    out = []

    def flush(b):
        m = pending.pop(b, None)
        if m is not None:
            out.append(EOp('mat1', (b,), m))

    for op in ops:
        if op.kind == 'mat1':
            b = op.bits[0]
            pending[b] = op.mat @ pending[b] if b in pending else op.mat.copy()
            continue
        if op.kind == 'norm':
            for b in sorted(pending):
                flush(b)
        else:
            for b in op.bits:
                flush(b)
        out.append(op)
    for b in sorted(pending):
        flush(b)
    return out


_OP2_MAT = {'cx': _CX_TC, 'cz': _CZ}


def fuse_2q(ops):
    """Fold dense 1-qubit ops into a neighbouring 2-qubit op.

    For each 2-qubit op (mat2 / cx / cz) in order, the candidates are the mat1
    that is the last op on either of its bits before it and the mat1 that is
    the next op on either bit after it (nothing else touches that bit in
    between, so moving the 1-qubit gate to the 2-qubit op keeps the product).
    The ops are replaced by one mat2 (post_a⊗post_b)·G·(pre_a⊗pre_b) when that
    saves ALU work: always for a dense mat2, and for a CNOT/CZ (nearly free
    as a permutation / sign) when it absorbs at least three 1-qubit gates, the
    point where one 16-FMA 2-qubit product beats the 8-FMA 1-qubit products.
    Ops are assumed 1-qubit-fused already (at most one mat1 per bit between
    other ops)."""
    N = len(ops)
    allbits = set()
    for op in ops:
        allbits.update(op.bits)
    prev = [dict() for _ in range(N)]
    nxt = [dict() for _ in range(N)]
    last = {}
    for i, op in enumerate(ops):
        touched = allbits if op.kind == 'norm' else op.bits
        for b in touched:
            p = last.get(b)
            prev[i][b] = p
            if p is not None:
                nxt[p][b] = i
            last[b] = i
    dead = set()
    repl = {}
    for j, op in enumerate(ops):
        if op.kind not in ('mat2', 'cx', 'cz'):
            continue
        a, b = op.bits
        pre, post = {}, {}
        for bit in (a, b):
            p = prev[j].get(bit)
            if p is not None and p not in dead and ops[p].kind == 'mat1':
                pre[bit] = p
            q = nxt[j].get(bit)
            if q is not None and q not in dead and ops[q].kind == 'mat1':
                post[bit] = q
        k = len(pre) + len(post)
        if 8.0 * k + OP_COST[op.kind] - OP_COST['mat2'] < 8.0:
            continue
        eye = np.eye(2, dtype=np.complex128)

        def side(d):
            return np.kron(ops[d[a]].mat if a in d else eye, ops[d[b]].mat if b in d else eye)
        g = op.mat if op.kind == 'mat2' else _OP2_MAT[op.kind].astype(np.complex128)
        repl[j] = EOp('mat2', (a, b), side(post) @ g @ side(pre))
        dead.update(pre.values())
        dead.update(post.values())
    return [repl.get(i, op) for i, op in enumerate(ops) if i not in dead]


# ----------------------------------------------------------------- scheduling
@dataclass
class Stage:
    rbits: list
    ops: list = field(default_factory=list)


@dataclass
class Pass:
    kind: int
    bits: list                 # global bits (tile: ascending tile-local order; dense: matrix order)
    ops: list = field(default_factory=list)
    tb: int = 0
    stages: list = field(default_factory=list)
    mat: np.ndarray = None
    flags: int = 0
    rb: int = RB
    # remapped store (tile-local bit x is written to physical bit store_bits[x]);
    # the last layout keeps store_lanes off the registers and store_regs on them
    store_bits: list = None
    store_lanes: list = None
    store_regs: list = None


@dataclass
class TileConfig:
    """Launch shape of tile passes (picked by the autotuner).

    tb: log2 threads per workgroup (tile L = tb + 4 bits);
    low_bits: lowest address bits every tile contains (contiguous runs of
              2^low_bits amplitudes per load);
    max_cost: cap on Σ per-amplitude VALU cost of the ops in one pass;
    prefetch: persistent workgroups that copy the next tile HBM -> LDS while
              computing the current one (forward passes only);
    specialize: compile each forward pass into its own straight-line kernel
              (qaravan_amd.specialize) instead of the op-interpreting kernel;
    fuse2q: fold 1-qubit gates into neighbouring 2-qubit gates (fuse_2q);
    io_lane_bits: tile-local bits that the first and last stage of a pass keep
              off the registers, so lanes 0..2^io_lane_bits-1 load and store
              consecutive amplitudes (4 = 128-byte runs for complex64);
    rb: register bits per thread (tile L = tb + rb): 4 everywhere; 5 in the
              per-pass compiled forward kernels (32 amplitudes per thread,
              fewer LDS re-distributions per pass); 3 in the per-pass compiled
              forward and adjoint kernels (8 amplitudes per thread: twice the
              waves for small states).
    lds_split: per-pass compiled kernels move the tile (and λ) through LDS
              in 2^lds_split parts (0, 1 or 2), split on register bits that
              the two layouts of a re-distribution share: 1/2^lds_split of the
              LDS per workgroup, so more workgroups per CU (the stage planner
              keeps that many shared bits).
    remap: per-pass compiled kernels store each tile with its bits
              permuted, so that the qubits the next passes need most come to
              the low address bits every tile holds; fewer passes, plus one
              permutation back to the natural order at the end (none for an
              adjoint sweep, whose ψ and λ are scratch afterwards).
    remap_free: a remapped pass takes the lanes its last layout already has
              when they hold at least this many of the qubits wanted next
              (no extra layout change); otherwise it adds one layout change
              to bring the wanted qubits to the lanes (-1: always).
    generic:  tile passes run in the generic kernels (one compiled kernel
              for every pass structure; stages dispatched to the pattern
              library or interpreted): no compiler at run time.  3 or 4
              register bits, tb >= 6, no LDS split; remap allowed.
    """
    tb: int = 8
    low_bits: int = 6
    max_cost: float = 96.0
    prefetch: bool = False
    specialize: bool = False
    fuse2q: bool = True
    io_lane_bits: int = 6
    rb: int = RB
    lds_split: int = 0
    remap: bool = False
    remap_free: int = -1
    generic: bool = False


def _is_barrier(op):
    return op.kind in ('dense', 'norm')


def is_diagonal(op):
    """Diagonal in the computational basis: such ops commute with each other,
    so a diagonal op may move past another diagonal op on shared bits.
    (Cached on the op: the schedulers ask about every op many times.)"""
    d = op.__dict__.get('_diag')
    if d is None:
        d = op.__dict__['_diag'] = _is_diagonal(op)
    return d


def _is_diagonal(op):
    if op.kind == 'cz':
        return True
    if op.kind in ('rot1', 'rot2'):
        return op.pauli == PAULI_CODE['z']
    if op.kind in ('mat1', 'mat2'):
        m = np.asarray(op.mat)
        return not np.any(m - np.diag(np.diag(m)))
    return False


class _Blocker:
    """Bits held back by ops that were skipped: a later op must wait if it
    shares a bit with a skipped op, unless both are diagonal (they commute)."""

    def __init__(self):
        self.diag, self.other = set(), set()

    def blocks(self, op, bits):
        return bool(bits & self.other) if is_diagonal(op) else bool(bits & (self.other | self.diag))

    def add(self, op, bits):
        (self.diag if is_diagonal(op) else self.other).update(bits)

    def all(self):
        return self.diag | self.other


def partition(ops, n, cfg):
    """Greedy packing of ops into passes (see module docstring, step 3)."""
    L = cfg.tb + cfg.rb
    if L > n:
        raise ValueError(f"tile of {L} bits exceeds n={n}")
    m = min(cfg.low_bits, L)
    low = set(range(m))
    rem = list(ops)
    passes = []
    while rem:
        head = rem[0]
        if head.kind == 'dense':
            passes.append(Pass(PASS_DENSE, list(head.bits), [head], mat=head.mat))
            rem.pop(0)
            continue
        if head.kind == 'norm':
            passes.append(Pass(PASS_NORMALIZE, [], [head]))
            rem.pop(0)
            continue
        T = set(low)
        blocked = _Blocker()
        taken = []
        cost = 0.0
        for i, op in enumerate(rem):
            if _is_barrier(op):
                break
            b = set(op.bits)
            if blocked.blocks(op, b):
                blocked.add(op, b)
            else:
                need = b - T
                c = OP_COST[op.kind]
                if len(T) + len(need) <= L and (cost + c <= cfg.max_cost or not taken):
                    T |= need
                    taken.append(i)
                    cost += c
                else:
                    blocked.add(op, b)
            if len(blocked.other) == n or cost + 0.25 > cfg.max_cost:
                break   # nothing later can join this pass
        # complete the tile with the lowest unused bits
        x = 0
        while len(T) < L:
            if x not in T:
                T.add(x)
            x += 1
        tset = set(taken)
        p = Pass(PASS_TILE, sorted(T), [rem[i] for i in taken], tb=cfg.tb,
                 flags=PASSF_PREFETCH if cfg.prefetch else 0, rb=cfg.rb)
        passes.append(p)
        rem = [op for i, op in enumerate(rem) if i not in tset]
    return passes


def partition_remap(ops, n, cfg, pos0=None):
    """Greedy packing like ``partition``, with qubits moving between address
    bits.  ``pos[q]`` is the physical address bit of logical bit q; ops are
    placed in passes with physical bits.  After each pass the store permutes
    the tile's bits so that up to ``low_bits`` qubits that the remaining ops
    need soonest move to the low physical bits (which every tile holds); the
    qubits they displace take the vacated positions.  So that the store stays
    coalesced, the moved qubits come from the tile's low + rb lowest bits:
    then the last layout can hold every lower non-moving tile bit in
    registers and put the moved ones on lanes 0..low-1 in order.  Returns
    (passes, pos) with pos the final layout.  ``pos0``: the layout the
    program starts from (default: the natural order)."""
    L = cfg.tb + cfg.rb
    low = min(cfg.low_bits, L)
    if L > n:
        raise ValueError(f"tile of {L} bits exceeds n={n}")
    pos = list(pos0) if pos0 is not None else list(range(n))
    rem = list(ops)
    passes = []
    while rem:
        if _is_barrier(rem[0]):
            raise ValueError("remap takes circuits of tile ops only (no dense or reset passes)")
        T = set(range(low))
        blocked = _Blocker()
        taken = []
        cost = 0.0
        for i, op in enumerate(rem):
            if _is_barrier(op):
                break
            b = set(op.bits)                           # logical: ordering
            if blocked.blocks(op, b):
                blocked.add(op, b)
            else:
                need = {pos[x] for x in b} - T         # physical: tile
                c = OP_COST[op.kind]
                if len(T) + len(need) <= L and (cost + c <= cfg.max_cost or not taken):
                    T |= need
                    taken.append(i)
                    cost += c
                else:
                    blocked.add(op, b)
            if len(blocked.other) == n or cost + 0.25 > cfg.max_cost:
                break
        x = 0
        while len(T) < L:
            if x not in T:
                T.add(x)
            x += 1
        tset = set(taken)
        pops = [EOp(o.kind, tuple(pos[x] for x in o.bits), o.mat, o.param, o.pauli, o.flags)
                for o in (rem[i] for i in taken)]
        rem = [op for i, op in enumerate(rem) if i not in tset]
        Ts = sorted(T)
        inv = {ph: q for q, ph in enumerate(pos)}
        # qubits wanted on the low bits: those the remaining ops need soonest
        wanted = []
        for op in rem:
            if _is_barrier(op) or len(wanted) >= low + 2:
                break
            for q in op.bits:
                if q not in wanted:
                    wanted.append(q)
        p = Pass(PASS_TILE, Ts, pops, tb=cfg.tb, rb=cfg.rb)
        pos_in = list(pos)
        _stage_ops(p, cfg.io_lane_bits, keep_shared=cfg.lds_split, free_store=True)
        good = {Ts.index(pos[q]) for q in wanted if pos[q] in T}
        lanes = _remap_last_stage(p, good, cfg.lds_split, low)
        # lanes the last layout gives for free, or the soonest-needed qubits of
        # the tile's low + rb lowest bits behind one more layout change
        top = [Ts.index(pos[q]) for q in wanted[:low] if pos[q] in T and Ts.index(pos[q]) < low + cfg.rb]
        if cfg.remap_free < 0 or sum(x in lanes for x in top) < min(len(top), cfg.remap_free):
            cand = {inv[ph] for ph in Ts[:low + cfg.rb]}
            new_low = []
            for op in rem:
                if _is_barrier(op) or len(new_low) >= low:
                    break
                new_low += [q for q in op.bits if q in cand and q not in new_low]
            new_low = new_low[:low]
            for q in [inv[ph] for ph in range(low)]:
                if len(new_low) < low and q not in new_low:
                    new_low.append(q)
            lanes = sorted(Ts.index(pos[q]) for q in new_low)
            p = Pass(PASS_TILE, Ts, pops, tb=cfg.tb, rb=cfg.rb)
            p.store_lanes = lanes
            _stage_ops(p, cfg.io_lane_bits, keep_shared=cfg.lds_split)
        dest = {x: ph for x, ph in enumerate(Ts)}
        vacated = sorted(Ts[x] for x in lanes if Ts[x] >= low)
        displaced = sorted(x for x in range(low) if x not in lanes)
        for x, ph in zip(displaced, vacated):
            dest[x] = ph
        for i, x in enumerate(lanes):
            dest[x] = i
        p.pos_in = pos_in             # the layout this pass starts from
        if any(dest[x] != Ts[x] for x in range(L)):
            p.store_bits = [dest[x] for x in range(L)]
            p.store_lanes = lanes
            p.store_regs = [x for x in range(lanes[-1]) if x not in lanes]
            p.flags |= PASSF_REMAP_STORE
            newpos = list(pos)
            for x in range(L):
                newpos[inv[Ts[x]]] = dest[x]
            pos = newpos
        passes.append(p)
    return passes, pos


def _stage_ops(p, io_lane_bits=6, keep_shared=0, free_store=False):
    """Split a tile pass's ops into register stages (tile-local bits).
    keep_shared: every stage keeps at least this many register bits of the
    stage before it (split LDS exchanges); the ops of a stage leave room.
    free_store: the store may remap bits (the lanes of the last layout go to
    the low address bits), so the last layout needs no extra stage."""
    RB = p.rb
    L = p.tb + RB
    loc = {g: i for i, g in enumerate(p.bits)}
    pend = [(op, tuple(loc[b] for b in op.bits)) for op in p.ops]
    lane_bits = set(range(min(io_lane_bits, L)))
    io_bits = list(range(L - RB, L))            # register bits of a coalesced layout
    stages = []
    cur = None
    while pend:
        if cur is not None:
            took, blocked, keep = [], _Blocker(), []
            creg = set(cur.rbits)
            for i, (op, lb) in enumerate(pend):
                s = set(lb)
                if not blocked.blocks(op, s) and s <= creg:
                    took.append((op, lb))
                else:
                    blocked.add(op, s)
                    keep.append((op, lb))
                    if creg <= blocked.other:       # every register bit is held back:
                        keep.extend(pend[i + 1:])   # nothing later can join this stage
                        break
            if took:
                cur.ops.extend(took)
                pend = keep
                continue
        def pick(avoid):
            R, blocked = [], _Blocker()
            for op, lb in pend:
                s = set(lb)
                if blocked.blocks(op, s) or s & avoid:
                    blocked.add(op, s)
                    if len(blocked.other) == L:     # every tile bit is held back
                        break
                    continue
                new = [x for x in lb if x not in R]
                # keep_shared: the last free slots are for bits of the previous layout
                room = RB - (max(0, keep_shared - len(set(R + new) & set(cur.rbits))) if cur is not None else 0)
                if len(R) + len(new) <= room:
                    R += new
                else:
                    blocked.add(op, s)
                if len(R) == RB:
                    break
            return R
        # the first layout is also the load layout: prefer ops off the lane bits
        R = pick(lane_bits) if cur is None else []
        if not R:
            R = pick(set())
        if cur is not None:                     # bits both layouts hold
            for x in sorted(cur.rbits, reverse=True):
                if len(set(R) & set(cur.rbits)) >= keep_shared or len(R) == RB:
                    break
                if x not in R:
                    R.append(x)
        for x in reversed(range(L)):            # fill with high tile bits
            if len(R) == RB:
                break
            if x not in R:
                R.append(x)
        cur = Stage(R, [])
        stages.append(cur)
    if not stages:
        stages.append(Stage(list(io_bits), []))
    # coalesced load/store: the first/last layout must keep lanes on low bits
    if set(stages[0].rbits) & lane_bits:
        stages.insert(0, Stage(_io_layout(stages[0].rbits, io_bits, lane_bits, keep_shared), []))
    if p.store_lanes is not None:
        # remapped store: the lanes carry the tile bits bound for the low
        # physical bits, in order, so every lower tile bit sits in a register
        sl = list(p.store_lanes[:min(io_lane_bits, len(p.store_lanes))])
        req = [x for x in range(sl[-1]) if x not in sl]
        last = stages[-1]
        if set(last.rbits) & set(sl) or not set(req) <= set(last.rbits):
            R = list(req)
            for x in sorted(last.rbits, reverse=True) + list(range(L - 1, -1, -1)):
                if len(R) == RB:
                    break
                if x not in R and x not in sl:
                    R.append(x)
            stages.append(Stage(R, []))
    elif set(stages[-1].rbits) & lane_bits and not free_store:
        stages.append(Stage(_io_layout(stages[-1].rbits, io_bits, lane_bits, keep_shared), []))
    p.stages = stages
    return p


def _remap_last_stage(p, good, keep_shared, low):
    """Register bits of a remapped pass's last layout, and its lanes (the
    ``low`` lowest non-register tile bits, which the store sends to address
    bits 0..low-1 in order).  The bits the last ops use stay in registers; the
    free register slots first keep ``keep_shared`` bits of the layout before
    (split LDS exchanges), then take low tile bits whose qubits are not in
    ``good`` (wanted soon), so that wanted qubits become lanes; the rest are
    high tile bits."""
    last = p.stages[-1]
    RB = p.rb
    L = p.tb + RB
    used = {x for _, lb in last.ops for x in lb}
    prev = set(p.stages[-2].rbits) if len(p.stages) > 1 else set()
    R = set(used)
    # shared bits for the split exchange: prefer unwanted low ones
    need = max(0, keep_shared - len(R & prev)) if prev else 0
    for x in sorted(prev - R, key=lambda x: (x in good, -x if x >= low + RB else x)):
        if need == 0 or len(R) == RB:
            break
        R.add(x)
        need -= 1
    # push unwanted low bits into registers so wanted ones become lanes
    lanes = []
    for x in range(L):
        if len(lanes) == low:
            break
        if x in R:
            continue
        if x not in good and len(R) < RB and any(g > x and g not in R for g in good):
            R.add(x)
        else:
            lanes.append(x)
    for x in range(L - 1, -1, -1):                  # fill above the lanes
        if len(R) == RB:
            break
        if x not in R and x not in lanes and x > lanes[-1]:
            R.add(x)
    for x in range(L - 1, -1, -1):                  # (cannot happen unless L is tiny)
        if len(R) == RB:
            break
        if x not in R and x not in lanes:
            R.add(x)
    ordered = [x for x in last.rbits if x in R] + sorted(R - set(last.rbits))
    p.stages[-1] = Stage(ordered[:RB], last.ops)
    nonreg = [x for x in range(L) if x not in R]
    return nonreg[:low]


def _io_layout(nb, io_bits, lane_bits, keep_shared):
    """Register bits of a coalesced first/last layout next to the stage
    holding ``nb``: the high tile bits, or (keep_shared) the neighbour's
    non-lane register bits topped up with high tile bits."""
    if not keep_shared:
        return list(io_bits)
    R = [x for x in nb if x not in lane_bits]
    for x in range(max(io_bits), -1, -1):
        if len(R) == len(io_bits):
            break
        if x not in R and x not in lane_bits:
            R.append(x)
    return R


# ----------------------------------------------------------------- encoding
def _swap_qubits_4x4(m):
    return m.reshape(2, 2, 2, 2).transpose(1, 0, 3, 2).reshape(4, 4)


class Program:
    """An encoded program plus the metadata the host needs."""

    def __init__(self, n, passes, blob, grad_params, n_params):
        self.n = n
        self.passes = passes
        self.blob = blob                  # np.uint8 array, the svh program
        self.grad_params = grad_params    # slot -> param index (all passes)
        self.n_params = n_params
        self._dev = {}
        self._sp = {}
        self.specialize = False
        self.final_pos = None             # remapped programs: physical bit of logical bit q at the end
        self.final_cx = []                # (control, target) physical bits of CNOTs the final permutation applies
        self.end_pos = None               # keep_layout programs: the layout the state is left in (None: natural)

    def specialized(self, dtype_code, dual=False):
        """Loaded straight-line kernels of this program (compiled on first use)."""
        key = (dtype_code, dual)
        if key not in self._sp:
            from .specialize import specialized_for
            self._sp[key] = specialized_for(self, dtype_code, dual)
        return self._sp[key]

    @property
    def n_passes(self):
        return len(self.passes)

    def n_tile_passes(self):
        return sum(p.kind == PASS_TILE for p in self.passes)

    def device_copy(self, device):
        import torch
        key = str(device)
        if key not in self._dev:
            self._dev[key] = torch.from_numpy(self.blob).to(device)
        return self._dev[key]

    def describe(self):
        kinds = {PASS_TILE: 0, PASS_DENSE: 0, PASS_NORMALIZE: 0}
        st = 0
        for p in self.passes:
            kinds[p.kind] += 1
            st += len(p.stages)
        return dict(passes=len(self.passes), tile=kinds[PASS_TILE], dense=kinds[PASS_DENSE],
                    normalize=kinds[PASS_NORMALIZE], stages=st,
                    ops=sum(len(p.ops) for p in self.passes))


def encode(n, passes, grad=False, n_params=0, final_pos=None, final_cx=()):
    stage_rows, op_rows, grad_params, pool = [], [], [], []
    pass_rows = np.zeros(len(passes), dtype=PASS_DT)
    for pi, p in enumerate(passes):
        pr = pass_rows[pi]
        pr['kind'] = p.kind
        pr['rb'] = p.rb
        pr['flags'] = p.flags
        if p.kind == PASS_TILE:
            L = p.tb + p.rb
            pr['L'] = L
            pr['tb'] = p.tb
            pr['bits'][:L] = p.bits
            pr['stage_begin'] = len(stage_rows)
            pr['n_stages'] = len(p.stages)
            pr['grad_begin'] = len(grad_params)
            slot = 0
            for st in p.stages:
                s = np.zeros((), dtype=STAGE_DT)
                s['pattern'] = -1
                s['rbits'][:p.rb] = st.rbits
                tab = stage_tables(p.bits, p.tb, st.rbits)
                s['rsorted'][:p.rb] = tab['rsorted']
                s['tgbit'][:p.tb] = tab['tgbit']
                s['lsw'][:1 << p.rb] = tab['lsw']
                s['goff'][:1 << p.rb] = tab['goff']
                s['op_begin'] = len(op_rows)
                ridx = {x: k for k, x in enumerate(st.rbits)}
                for op, lb in st.ops:
                    regs = [ridx[x] for x in lb]
                    m = _ZERO_M
                    a = regs[0]
                    b = 0
                    if op.kind == 'mat1':
                        kind = OP_MAT1
                        m = np.zeros(32)
                        m[:8] = np.ascontiguousarray(op.mat, dtype=np.complex128).reshape(-1).view(np.float64)
                    elif op.kind == 'mat2':
                        mm = op.mat
                        a, b = regs
                        if a < b:
                            a, b = b, a
                            mm = _swap_qubits_4x4(mm)
                        kind = OP_MAT2
                        m = np.ascontiguousarray(mm, dtype=np.complex128).reshape(-1).view(np.float64)
                    elif op.kind == 'cx':
                        kind, a, b = OP_CX, regs[0], regs[1]
                    elif op.kind == 'cz':
                        kind, a, b = OP_CZ, max(regs), min(regs)
                    elif op.kind == 'rot1':
                        kind = OP_ROT1
                    elif op.kind == 'rot2':
                        kind, a, b = OP_ROT2, max(regs), min(regs)
                    else:
                        raise ValueError(op.kind)
                    gslot = 0
                    if grad and op.kind in ('rot1', 'rot2') and (op.flags & OPF_GRAD):
                        gslot = slot
                        grad_params.append(op.param)
                        slot += 1
                    op_rows.append((kind, a, b, op.param, op.pauli, op.flags, gslot, 0, m, m.astype(np.float32)))
                s['op_end'] = len(op_rows)
                stage_rows.append(s)
            if p.store_bits is not None:
                # the store record: global bits of the last layout after the remap
                s = np.zeros((), dtype=STAGE_DT)
                s['pattern'] = -1
                last = p.stages[-1].rbits
                tab = stage_tables(p.store_bits, p.tb, last)
                s['rbits'][:p.rb] = last
                s['rsorted'][:p.rb] = tab['rsorted']
                s['tgbit'][:p.tb] = tab['tgbit']
                s['lsw'][:1 << p.rb] = tab['lsw']
                s['goff'][:1 << p.rb] = tab['goff']
                s['op_begin'] = s['op_end'] = len(op_rows)
                stage_rows.append(s)
            pr['n_grad'] = slot
        elif p.kind == PASS_DENSE:
            k = len(p.bits)
            pr['L'] = k
            pr['bits'][:k] = p.bits
            pr['pool_off'] = len(pool)
            pool.extend(np.ascontiguousarray(p.mat, dtype=np.complex128).reshape(-1).view(np.float64))
    stages = np.array(stage_rows, dtype=STAGE_DT) if stage_rows else np.zeros(0, STAGE_DT)
    opsa = np.array(op_rows, dtype=OP_DT) if op_rows else np.zeros(0, OP_DT)
    for pi, p in enumerate(passes):
        if p.kind == PASS_TILE and (p.flags & PASSF_GENERIC):
            # stages the generic kernels run from the compiled pattern library
            from .patterns import pattern_ids
            s0, ns = int(pass_rows[pi]['stage_begin']), int(pass_rows[pi]['n_stages'])
            stages['pattern'][s0:s0 + ns] = pattern_ids(stages[s0:s0 + ns], opsa, p.rb, grad)
    used = [int(r[3]) + 1 for r in op_rows if r[0] in (OP_ROT1, OP_ROT2)]
    grads = np.array(grad_params, dtype=np.int32)
    poola = np.array(pool, dtype=np.float64)
    perm = (np.array(final_pos, dtype=np.int32) if final_pos is not None and list(final_pos) != list(range(n))
            else np.zeros(0, np.int32))
    if final_cx and not perm.size:
        raise ValueError("final CNOTs need a final permutation")
    if len(final_cx) > MAX_FINAL_CX:
        raise ValueError(f"at most {MAX_FINAL_CX} final CNOTs")
    # int32 count, then (control, target) pairs
    cxa = (np.array([len(final_cx)] + [b for cx in final_cx for b in cx], dtype=np.int32) if final_cx
           else np.zeros(0, np.int32))

    def align(x):
        return (x + 63) & ~63
    off = align(HDR_DT.itemsize)
    offs = {}
    for name, arr in (('passes', pass_rows), ('stages', stages), ('ops', opsa), ('grad', grads),
                      ('pool', poola), ('perm', perm), ('cx', cxa)):
        offs[name] = off
        off = align(off + arr.nbytes)
    blob = np.zeros(off, dtype=np.uint8)
    h = np.zeros((), dtype=HDR_DT)
    h['magic'], h['version'], h['n_qubits'] = MAGIC, ABI_VERSION, n
    h['n_passes'], h['n_stages'], h['n_ops'] = len(pass_rows), len(stages), len(opsa)
    h['n_grad_slots'], h['n_pool'] = len(grads), len(poola)
    h['n_params'] = max([int(n_params)] + used) if used else 0
    for name in offs:
        h['off_' + name] = offs[name]
    if not perm.size:
        h['off_perm'] = 0
    if not cxa.size:
        h['off_cx'] = 0
    blob[:HDR_DT.itemsize] = np.frombuffer(h.tobytes(), dtype=np.uint8)
    for name, arr in (('passes', pass_rows), ('stages', stages), ('ops', opsa), ('grad', grads),
                      ('pool', poola), ('perm', perm), ('cx', cxa)):
        if arr.nbytes:
            blob[offs[name]:offs[name] + arr.nbytes] = np.frombuffer(arr.tobytes(), dtype=np.uint8)
    prog = Program(n, passes, blob, grads, n_params)
    prog.final_pos = list(final_pos) if perm.size else None
    prog.final_cx = [tuple(int(b) for b in cx) for cx in final_cx]
    return prog


def min_tile_n(grad=False):
    return 8


def compile_ops(ops, n, cfg=None, grad=False, n_params=0, fuse=True, keep_layout=False, pos0=None):
    """EOps -> Program.  Small registers (n < 8) use one dense pass per op.

    keep_layout: a remapped program ends in its own layout (``prog.end_pos``:
    physical bit of logical bit q) instead of permuting back to the natural
    order.  pos0: the layout the input state is in (physical bit of logical
    bit q; default natural), e.g. the end_pos of the program that made it."""
    cfg = cfg or TileConfig()
    if pos0 is not None:
        pos0 = [int(x) for x in pos0]
        if sorted(pos0) != list(range(n)):
            raise ValueError("pos0 is not a permutation of the address bits")
        if pos0 == list(range(n)):
            pos0 = None
        elif not (grad or keep_layout):
            # nothing would bring such a state back to the natural order
            raise ValueError("a permuted input layout (pos0) needs keep_layout or an adjoint program")
    if fuse:
        ops = fuse_1q(ops)
        if cfg.fuse2q:
            ops = fuse_2q(ops)
    if n < min_tile_n():
        if pos0 is not None:
            raise NotImplementedError("a permuted input layout needs n >= 8")
        if any(op.kind in ('rot1', 'rot2') for op in ops):
            raise NotImplementedError("device-angle rotations need n >= 8")
        passes = []
        for op in ops:
            if op.kind == 'norm':
                passes.append(Pass(PASS_NORMALIZE, [], [op]))
            else:
                passes.append(Pass(PASS_DENSE, list(_dense_bits(op)), [op], mat=_dense_mat(op)))
        return encode(n, passes, grad, n_params)
    if cfg.rb not in (3, 4, 5):
        raise ValueError("rb (register bits per thread) must be 3, 4 or 5")
    if cfg.generic and (cfg.specialize or cfg.prefetch or cfg.lds_split or cfg.rb not in (3, 4)):
        raise ValueError("generic passes take 3 or 4 register bits, no LDS split, and exclude the compiled "
                         "per-pass and prefetching kernels")
    if cfg.rb == 5 and (grad or cfg.prefetch or not cfg.specialize):
        raise ValueError("rb=5 runs only in the per-pass compiled forward kernels (specialize=True, "
                         "no prefetch, no adjoint)")
    if cfg.rb == 3 and (cfg.prefetch or not (cfg.specialize or cfg.generic)):
        raise ValueError("rb=3 runs only in the per-pass compiled or the generic kernels (no prefetch)")
    if cfg.lds_split and (cfg.prefetch or not cfg.specialize):
        raise ValueError("lds_split runs only in the per-pass compiled kernels (specialize=True, no prefetch)")
    if cfg.lds_split not in (0, 1, 2) or cfg.lds_split >= cfg.rb:
        raise ValueError("lds_split must be 0, 1 or 2 and below the register bits")
    tb = min(cfg.tb, n - cfg.rb)
    if cfg.generic and tb < 6:
        raise ValueError("generic passes need tb >= 6 (whole waves): n >= rb + 6")
    split = cfg.lds_split
    remap = (cfg.remap and (cfg.specialize or cfg.generic) and not cfg.prefetch
             and not any(_is_barrier(op) for op in ops))
    cfg = TileConfig(tb=tb, low_bits=max(1, cfg.low_bits), max_cost=cfg.max_cost,
                     prefetch=cfg.prefetch and not grad, specialize=cfg.specialize, fuse2q=cfg.fuse2q,
                     io_lane_bits=cfg.io_lane_bits, rb=cfg.rb, lds_split=split, remap=remap,
                     remap_free=cfg.remap_free, generic=cfg.generic)
    final_pos = None
    final_cx = []
    end_pos = pos0
    if remap:
        passes, final_pos = partition_remap(ops, n, cfg, pos0)
        end_pos = final_pos
        if grad or keep_layout:
            # an adjoint sweep only accumulates the gradient: the ψ and λ it
            # leaves behind are scratch; a program that keeps its layout hands
            # it to the next one: no permutation back at the end
            final_pos = None
        else:
            passes, final_pos, final_cx = _fold_final_cx(passes, final_pos, n)
    else:
        if pos0 is not None:
            # no remapping: the ops act on the physical bits of the input layout
            ops = [EOp(o.kind, tuple(pos0[b] for b in o.bits), o.mat, o.param, o.pauli, o.flags) for o in ops]
        passes = partition(ops, n, cfg)
    for p in passes:
        if p.kind == PASS_TILE:
            if cfg.generic:
                p.flags |= PASSF_GENERIC
            if not remap:
                _stage_ops(p, cfg.io_lane_bits, keep_shared=split)
            shared = min((len(set(a.rbits) & set(b.rbits)) for a, b in zip(p.stages, p.stages[1:])), default=RB)
            if split >= 2 and shared >= 2:
                p.flags |= PASSF_QUARTER_LDS
            elif split >= 1 and shared >= 1:
                p.flags |= PASSF_HALF_LDS
    prog = encode(n, passes, grad, n_params, final_pos, final_cx)
    prog.specialize = bool(cfg.specialize)
    if keep_layout and end_pos is not None and list(end_pos) != list(range(n)):
        prog.end_pos = list(end_pos)
    return prog


def zero_start_variant(prog, dtype_code):
    """The same program with its first pass generating |0…0⟩ in registers
    instead of reading the state (the caller then skips the initialisation
    sweep), or None when that pass does not run in a compiled kernel of its
    own (the only kernels that generate)."""
    from .specialize import eligible
    if not getattr(prog, "specialize", False) or not prog.passes or prog.passes[0].kind != PASS_TILE:
        return None
    off = int(np.frombuffer(prog.blob[:HDR_DT.itemsize].tobytes(), dtype=HDR_DT)[0]['off_passes'])
    row = np.frombuffer(prog.blob[off:off + PASS_DT.itemsize].tobytes(), dtype=PASS_DT)[0].copy()
    if int(row['flags']) & (PASSF_GENERIC | PASSF_PREFETCH) or not eligible(row, dtype_code, False):
        return None
    row['flags'] = int(row['flags']) | PASSF_GEN_ZERO
    blob = prog.blob.copy()
    blob[off:off + PASS_DT.itemsize] = np.frombuffer(row.tobytes(), dtype=np.uint8)
    out = Program(prog.n, prog.passes, blob, prog.grad_params, prog.n_params)
    out.specialize = True
    out.final_pos, out.final_cx, out.end_pos = prog.final_pos, prog.final_cx, prog.end_pos
    return out


def _fold_final_cx(passes, final_pos, n):
    """A remapped program whose last tile pass holds only CNOTs (at most
    MAX_FINAL_CX): those CNOTs are linear maps of the address bits, so the
    final permutation applies them while it gathers (one sweep fewer).  The
    program then ends in the layout the dropped pass started from.  Returns
    (passes, final layout, [(control, target)] physical bits in gate order)."""
    last = passes[-1] if passes else None
    if (len(passes) < 2 or last.kind != PASS_TILE or not last.ops or len(last.ops) > MAX_FINAL_CX
            or any(o.kind != 'cx' for o in last.ops) or getattr(last, 'pos_in', None) is None
            or list(last.pos_in) == list(range(n))):
        return passes, final_pos, []
    # EOp cx bits are (target, control)
    return passes[:-1], list(last.pos_in), [(o.bits[1], o.bits[0]) for o in last.ops]


def _dense_bits(op):
    if op.kind == 'cx':
        return (op.bits[0], op.bits[1])
    return op.bits


def _dense_mat(op):
    if op.kind in ('mat1', 'mat2', 'dense'):
        return op.mat
    if op.kind == 'cx':
        return _CX_TC.astype(np.complex128)     # (target, control) order
    if op.kind == 'cz':
        return _CZ.astype(np.complex128)
    raise ValueError(op.kind)
