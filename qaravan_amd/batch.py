"""Batched forward + adjoint gradients for many parametrised circuits.

qaravan runs many small simulations one after another (e.g. the trajectories
of tensorQ/monte_carlo_sim.py:14-35, or one StatevectorSim per optimisation
step in applications/compilation.py:73-133).  Here a whole batch of circuits
of uneven qubit counts and depths, each with its own θ and Pauli-sum loss, is
one fixed launch sequence on one stream:

  init (all |0…0⟩) → trig tables → forward rounds → λ = Hψ and E → adjoint rounds

Round r launches pass r of every circuit that has one (one kernel per tile
shape), so small circuits share the GPU instead of each under-filling it.  The
sequence allocates nothing and reads nothing from the host, so it is captured
in one HIP graph and replayed as a single launch (``capture()``/``replay()``).
Per-circuit completion times come from the device clock (s_memrealtime,
100 MHz) written when a circuit's last gradient reduction finishes.
"""
from dataclasses import replace

import numpy as np
import torch

from . import _lib
from .adjoint import ParamProgram, term_table
from .program import HDR_DT, PASS_DT, PASS_TILE, PASSF_HALF_LDS, PASSF_QUARTER_LDS, RB, TileConfig

ITEM_DT = np.dtype([('psi', '<u8'), ('lam', '<u8'), ('pass', '<u8'), ('stages', '<u8'), ('ops', '<u8'),
                    ('trig', '<u8'), ('gpart', '<u8'), ('wg_begin', '<i8'), ('n', '<i4'), ('pad', '<i4')])
RED_DT = np.dtype([('part', '<u8'), ('slot_param', '<u8'), ('out', '<u8'), ('tstamp', '<u8'),
                   ('nblocks', '<i4'), ('ng', '<i4'), ('slot_begin', '<i8')])
HAM_DT = np.dtype([('psi', '<u8'), ('lam', '<u8'), ('terms', '<u8'), ('epart', '<u8'),
                   ('block_begin', '<i8'), ('n', '<i4'), ('nterms', '<i4')])
assert ITEM_DT.itemsize == 72 and RED_DT.itemsize == 48 and HAM_DT.itemsize == 48


def default_tb(n, dtype, rb=RB):
    """Workgroup size per circuit: a 12-bit tile (256 threads with 4 register
    bits, 512 with 3) once the state has room for it, otherwise the whole
    state in one tile."""
    return max(4, min(12 - rb, n - rb))


def batch_config(n, dtype, rb=3, remap_adjoint=True, threads=512):
    """(forward, adjoint) tile configurations of one circuit in the persistent
    batch: generic kernels, tiles of 2^12 amplitudes with 512-thread
    workgroups (2^11 with 256, and for complex128, so ψ and λ fit twice per
    CU's LDS) or the whole state when it is smaller, no pass cap.  The forward
    program keeps the natural bit order (the Hamiltonian step reads it); the
    adjoint sweep may remap, since its ψ and λ are scratch afterwards."""
    top = 12 if dtype == torch.complex64 else 11
    top = min(top, threads.bit_length() - 1 + rb)
    tb = max(6, min(top - rb, n - rb))
    cfg = TileConfig(tb=tb, low_bits=4, max_cost=1e9, io_lane_bits=4, rb=rb, generic=True)
    return cfg, replace(cfg, remap=bool(remap_adjoint))


def batch_programs(ns, circuits, op_lists, terms_list, n_params_list, dtype, tb_of_n, max_cost, keys, lds_split,
                   rb, remap):
    """The ParamProgram of every circuit of a round-synchronous batch (host
    only; circuits with equal ``keys`` share one).  Forward programs keep the
    natural bit order (the batched Hamiltonian reads ψ in it); adjoint sweeps
    may remap, their ψ and λ are scratch."""
    pps, cache = [], {}
    for c in range(len(ns)):
        key = None if keys is None else keys[c]
        if key is not None and key in cache:
            pps.append(cache[key])
            continue
        n = ns[c]
        cfg = TileConfig(tb=tb_of_n(n), low_bits=4, max_cost=max_cost, io_lane_bits=4,
                         specialize=bool(lds_split or remap) or rb != RB, lds_split=lds_split, rb=rb)
        adj_cfg = replace(cfg, remap=remap)
        if circuits is not None:
            pp = ParamProgram(circuits[c], terms_list[c], n_params_list[c], cfg=cfg, adj_cfg=adj_cfg, dtype=dtype)
        else:
            pp = ParamProgram(None, terms_list[c], n_params_list[c], cfg=cfg, adj_cfg=adj_cfg, dtype=dtype,
                              ops=op_lists[c], n=n)
        slots = pp.adj.grad_params.tolist()
        if len(set(slots)) != len(slots):
            # the round reductions add slot sums with atomics: a parameter
            # shared by several gates would get them in a varying order
            raise NotImplementedError("gates sharing a parameter: use persistent.PersistentBatch, which adds "
                                      "the slot sums of a parameter in slot order")
        if key is not None:
            cache[key] = pp
        pps.append(pp)
    return pps


def pass_grad_ops(p):
    """Gradient ops of one adjoint pass (its n_grad)."""
    return sum(1 for st in p.stages for op, _ in st.ops if op.flags & 2)


def gpart_rows(pp, n):
    """Doubles of gradient partials one circuit needs: per workgroup row, n_grad
    slots, for its largest adjoint pass."""
    m = 1
    for p in pp.adj.passes:
        if p.kind == PASS_TILE:
            m = max(m, (2 ** (n - p.tb - p.rb)) * max(pass_grad_ops(p), 1))
    return m


def spb_lds_bytes(tb, rb, split, dual, es, max_ngrad):
    """Dynamic LDS of one compiled batch-kernel launch: the tile (ψ, and λ in
    an adjoint pass) in 2^split parts, plus one row of gradient slots per wave."""
    L = tb + rb - split
    nw = max(1, 2 ** tb // 64)
    return (2 if dual else 1) * es * 2 ** L + (8 * nw * max_ngrad if dual else 0)


class BatchedAdjoint:
    """Energies and gradients of E_c(θ_c) = ⟨ψ_c(θ_c)|H_c|ψ_c(θ_c)⟩ for a list of
    circuits (rotations carry ``param_index`` into their own θ_c)."""

    def __init__(self, circuits, terms_list, n_params_list, device="cuda", dtype=torch.complex64,
                 tb_of_n=None, max_cost=1e9, keys=None, op_lists=None, ns=None, specialize=False, lds_split=0,
                 rb=RB, remap=False):
        """``circuits``: qaravan-style circuits (rotations carry ``param_index``),
        or None with ``op_lists`` (elementary ops) and ``ns``.  Circuits with equal
        ``keys`` share compiled programs (θ lives on the device, so programs
        depend on the structure only).  ``specialize``: run every pass with a
        kernel compiled for its structure (one launch per structure and round)
        instead of the op-interpreting batch kernel.  ``lds_split`` (with
        ``specialize``): the compiled kernels move tiles through LDS in
        2^lds_split parts, so more workgroups fit on a CU.  ``rb`` (with
        ``specialize``): register bits per thread, 3 or 4.  ``remap`` (with
        ``specialize``): the adjoint programs move qubits between address bits
        (fewer passes; the adjoint sweep needs no permutation back)."""
        _lib.require_gpu()
        if rb != RB:
            # 3 register bits passed the small-batch tests but the 512-circuit
            # benchmark batch (up to 22 qubits, 32 layers) faulted the GPU once,
            # cause not found: switched off until it is
            raise NotImplementedError("batched programs run with 4 register bits only")
        if (lds_split or remap) and not specialize:
            raise ValueError("lds_split and remap need the compiled batch kernels (specialize=True)")
        self.device = torch.device(device)
        self.dtype = dtype
        self.code = _lib.dtype_code(dtype)
        tb_of_n = tb_of_n or (lambda n: default_tb(n, dtype, rb))
        nc = len(circuits) if circuits is not None else len(op_lists)
        self.nc = nc
        self.ns = [c.num_sites for c in circuits] if circuits is not None else [int(n) for n in ns]
        if min(self.ns) < 10:
            raise NotImplementedError("batched Hamiltonian kernel works on blocks of 1024 amplitudes: n >= 10")
        self.pps = batch_programs(self.ns, circuits, op_lists, terms_list, n_params_list, dtype, tb_of_n, max_cost,
                                  keys, lds_split, rb, remap)
        es = 8 if dtype == torch.complex64 else 16
        # ---- flat device buffers
        self.amp_off = np.concatenate([[0], np.cumsum([2 ** n for n in self.ns])]).astype(np.int64)
        self.par_off = np.concatenate([[0], np.cumsum(n_params_list)]).astype(np.int64)
        total = int(self.amp_off[-1])
        dev = self.device
        self.psi = torch.empty(total, dtype=dtype, device=dev)
        self.lam = torch.empty(total, dtype=dtype, device=dev)
        self.theta = torch.zeros(int(self.par_off[-1]), dtype=torch.float64, device=dev)
        self.trig = torch.empty(int(self.par_off[-1]), dtype=dtype, device=dev)
        self.grad = torch.zeros(int(self.par_off[-1]), dtype=torch.float64, device=dev)
        self.energy = torch.zeros(nc, dtype=torch.float64, device=dev)
        self.t0 = torch.zeros(1, dtype=torch.int64, device=dev)
        self.t_end = torch.zeros(nc, dtype=torch.int64, device=dev)
        self.offsets = torch.tensor(self.amp_off[:-1], dtype=torch.int64, device=dev)
        self.cidx = torch.arange(nc, dtype=torch.int32, device=dev)
        # programs: one device blob holding every distinct forward and adjoint program
        blobs, where = [], {}
        off = 0
        for pp in self.pps:
            for prog in (pp.fwd, pp.adj):
                if id(prog) not in where:
                    where[id(prog)] = off
                    blobs.append((off, prog.blob))
                    off += (len(prog.blob) + 255) & ~255
        host = np.zeros(off, dtype=np.uint8)
        for o, b in blobs:
            host[o:o + len(b)] = b
        self.progs = torch.from_numpy(host).to(dev)
        pbase = self.progs.data_ptr()
        self.prog_addr = [(pbase + where[id(pp.fwd)], pbase + where[id(pp.adj)]) for pp in self.pps]
        # gradient partial buffers: the largest adjoint pass of each circuit
        gp_sizes = [gpart_rows(pp, self.ns[c]) for c, pp in enumerate(self.pps)]
        self.gp_off = np.concatenate([[0], np.cumsum(gp_sizes)]).astype(np.int64)
        self.gpart = torch.zeros(int(self.gp_off[-1]), dtype=torch.float64, device=dev)
        self.ep_off = np.concatenate([[0], np.cumsum([2 ** n // 1024 for n in self.ns])]).astype(np.int64)
        self.epart = torch.zeros(int(self.ep_off[-1]), dtype=torch.float64, device=dev)
        tt = [term_table(self.ns[c], terms_list[c]) for c in range(nc)]
        self.t_off = np.concatenate([[0], np.cumsum([len(t) for t in tt])]).astype(np.int64)
        self.terms = torch.from_numpy(np.concatenate(tt).view(np.uint8).copy()).to(dev)
        self.kernels = self._specialise() if specialize else None
        self._plan(es)
        self.graph = None

    # ---------------------------------------------------------------- plan
    def _specialise(self):
        """Kernel name of every (program, pass): passes with the same body text
        share one compiled batch kernel."""
        import hashlib

        from .specialize import BatchKernels, batch_body
        self.kname = {}
        kernels = {}
        for pp in self.pps:
            for dual, prog in ((False, pp.fwd), (True, pp.adj)):
                if id(prog) in self.kname:
                    continue
                names = []
                for pi, pa in enumerate(prog.passes):
                    body = batch_body(prog.blob, pi, self.code, dual)
                    nm = ("svh_spbd_" if dual else "svh_spb_") + hashlib.sha1(body.encode()).hexdigest()[:16]
                    kernels.setdefault(nm, (body, dual, pa.tb))
                    names.append(nm)
                self.kname[id(prog)] = names
        return BatchKernels(kernels, self.code)

    def _pass_ptrs(self, prog_addr, prog, pi):
        h = np.frombuffer(prog.blob[:HDR_DT.itemsize].tobytes(), dtype=HDR_DT)[0]
        return (prog_addr + int(h['off_passes']) + pi * PASS_DT.itemsize,
                prog_addr + int(h['off_stages']), prog_addr + int(h['off_ops']),
                prog_addr + int(h['off_grad']))

    def _plan(self, es):
        nc = self.nc
        psi0, lam0 = self.psi.data_ptr(), self.lam.data_ptr()
        trig0, gp0 = self.trig.data_ptr(), self.gpart.data_ptr()
        tables = []          # numpy arrays appended to one device blob
        self.launches = []   # (kind, args...) with table offsets resolved later

        def add_table(arr):
            off = sum((t.nbytes + 255) & ~255 for t in tables)
            tables.append(arr)
            return off

        def rounds(which):
            progs = [pp.fwd if which == 0 else pp.adj for pp in self.pps]
            R = max(len(p.passes) for p in progs)
            for r in range(R):
                groups = {}
                reds = []
                for c in range(nc):
                    prog = progs[c]
                    if r >= len(prog.passes):
                        continue
                    p = prog.passes[r]
                    tb = p.tb
                    L = tb + p.rb
                    n = self.ns[c]
                    ppass, pst, pops, pgrad = self._pass_ptrs(self.prog_addr[c][which], prog, r)
                    ng = pass_grad_ops(p) if which else 0
                    gkey = (tb, self.kname[id(prog)][r]) if self.kernels is not None else (tb, None)
                    split = 1 if p.flags & PASSF_HALF_LDS else 2 if p.flags & PASSF_QUARTER_LDS else 0
                    groups.setdefault(gkey, []).append(dict(split=split, rb=p.rb,
                        psi=psi0 + es * int(self.amp_off[c]), lam=(lam0 + es * int(self.amp_off[c])) if which else 0,
                        pass_=ppass, stages=pst, ops=pops, trig=trig0 + es * int(self.par_off[c]),
                        gpart=gp0 + 8 * int(self.gp_off[c]), nwg=2 ** (n - L), n=n, ng=ng))
                    if which and ng > 0:
                        blob = progs[c].blob
                        h = np.frombuffer(blob[:HDR_DT.itemsize].tobytes(), dtype=HDR_DT)[0]
                        pr = np.frombuffer(blob[int(h['off_passes']) + r * PASS_DT.itemsize:
                                                int(h['off_passes']) + (r + 1) * PASS_DT.itemsize].tobytes(),
                                           dtype=PASS_DT)[0]
                        last = r == len(prog.passes) - 1
                        reds.append(dict(part=gp0 + 8 * int(self.gp_off[c]),
                                         slot_param=pgrad + 4 * int(pr['grad_begin']),
                                         out=self.grad.data_ptr() + 8 * int(self.par_off[c]),
                                         tstamp=(self.t_end.data_ptr() + 8 * c) if last else 0,
                                         nblocks=2 ** (n - L), ng=ng))
                for (tb, kname), items in sorted(groups.items(), key=lambda kv: (kv[0][0], kv[0][1] or "")):
                    arr = np.zeros(len(items), dtype=ITEM_DT)
                    wg = 0
                    for i, it in enumerate(items):
                        arr[i] = (it['psi'], it['lam'], it['pass_'], it['stages'], it['ops'], it['trig'],
                                  it['gpart'], wg, it['n'], 0)
                        wg += it['nwg']
                    mng = max(it['ng'] for it in items)
                    if kname is None:
                        self.launches.append(('tile', add_table(arr), len(items), wg, tb, which, mng))
                    else:
                        # one kernel text per group, so one split for all its items
                        lds = spb_lds_bytes(tb, items[0]['rb'], items[0]['split'], bool(which), es, mng)
                        self.launches.append(('spb', add_table(arr), len(items), wg, tb, lds,
                                              self.kernels.index[kname]))
                if reds:
                    arr = np.zeros(len(reds), dtype=RED_DT)
                    sb = 0
                    for i, rd in enumerate(reds):
                        arr[i] = (rd['part'], rd['slot_param'], rd['out'], rd['tstamp'], rd['nblocks'], rd['ng'], sb)
                        sb += rd['ng']
                    self.launches.append(('reduce', add_table(arr), len(reds), sb))

        rounds(0)
        # λ = Hψ and energies
        ham = np.zeros(nc, dtype=HAM_DT)
        bb = 0
        tb0 = self.terms.data_ptr()
        for c in range(nc):
            n = self.ns[c]
            ham[c] = (psi0 + es * int(self.amp_off[c]), lam0 + es * int(self.amp_off[c]),
                      tb0 + 32 * int(self.t_off[c]), self.epart.data_ptr() + 8 * int(self.ep_off[c]), bb, n,
                      int(self.t_off[c + 1] - self.t_off[c]))
            bb += 2 ** n // 1024
        self.launches.append(('ham', add_table(ham), nc, bb))
        red = np.zeros(nc, dtype=RED_DT)
        for c in range(nc):
            red[c] = (self.epart.data_ptr() + 8 * int(self.ep_off[c]), self.cidx.data_ptr() + 4 * c,
                      self.energy.data_ptr(), 0, 2 ** self.ns[c] // 1024, 1, c)
        self.launches.append(('reduce', add_table(red), nc, nc))
        rounds(1)
        size = sum((t.nbytes + 255) & ~255 for t in tables)
        blob = np.zeros(max(size, 256), dtype=np.uint8)
        off = 0
        for t in tables:
            blob[off:off + t.nbytes] = np.frombuffer(t.tobytes(), dtype=np.uint8)
            off += (t.nbytes + 255) & ~255
        self.tables = torch.from_numpy(blob).to(self.device)
        self.n_launches = len(self.launches) + 3

    # ---------------------------------------------------------------- run
    def _launch_all(self):
        lib = _lib.load()
        st = _lib.stream_ptr(self.device)
        tb0 = self.tables.data_ptr()
        self.grad.zero_()
        self.energy.zero_()
        self.t_end.zero_()
        _lib.check(lib.svh_timestamp(_lib.ptr(self.t0), st), "svh_timestamp")
        _lib.check(lib.svh_batch_init(_lib.ptr(self.psi), self.psi.numel(), _lib.ptr(self.offsets), self.nc,
                                      self.code, st), "svh_batch_init")
        _lib.check(lib.svh_batch_trig(_lib.ptr(self.theta), self.theta.numel(), _lib.ptr(self.trig), self.code, st),
                   "svh_batch_trig")
        for L in self.launches:
            if L[0] == 'tile':
                _, off, ni, wg, tb, dual, mng = L
                _lib.check(lib.svh_batch_tile(tb0 + off, ni, wg, self.code, tb, dual, mng, st), "svh_batch_tile")
            elif L[0] == 'spb':
                _, off, ni, wg, tb, lds, k = L
                _lib.check(lib.svh_spb_launch(self.kernels.handle, k, tb0 + off, ni, wg, tb, lds, st),
                           "svh_spb_launch")
            elif L[0] == 'reduce':
                _, off, ni, nb = L
                _lib.check(lib.svh_batch_reduce(tb0 + off, ni, nb, st), "svh_batch_reduce")
            elif L[0] == 'ham':
                _, off, ni, nb = L
                _lib.check(lib.svh_batch_hamiltonian(tb0 + off, ni, nb, self.code, st), "svh_batch_hamiltonian")

    def set_theta(self, thetas):
        """θ for every circuit (list of arrays/tensors, or one flat tensor)."""
        if isinstance(thetas, torch.Tensor) and thetas.dim() == 1 and thetas.numel() == self.theta.numel():
            self.theta.copy_(thetas.to(torch.float64))
        else:
            flat = np.concatenate([np.asarray(t.detach().cpu() if hasattr(t, "detach") else t, dtype=np.float64)
                                   for t in thetas])
            self.theta.copy_(torch.from_numpy(flat))

    def run(self):
        """Eager run of the whole batch; returns (energies, flat gradient)."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch_all()
        return self.energy, self.grad

    def capture(self):
        """Record the whole batch into one HIP graph (after one eager warm-up
        run, so every kernel's one-time setup is done outside the capture)."""
        self._launch_all()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._launch_all()
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def gradients(self):
        g = self.grad.cpu().numpy()
        return [g[self.par_off[c]:self.par_off[c + 1]] for c in range(self.nc)]

    def latencies_us(self):
        """Per-circuit completion time since the batch started (device clock)."""
        t0 = int(self.t0.item())
        te = self.t_end.cpu().numpy().astype(np.int64)
        return (te - t0) / 100.0  # 100 MHz ticks -> µs


class BatchedForward:
    """Many fixed-matrix circuits on n qubits each, one state per circuit in a
    flat device buffer, run as pass-synchronous rounds: round r is pass r of
    every program, one svh_batch_tile launch per tile shape.  Used for
    Monte-Carlo trajectories (tensorQ/monte_carlo_sim.py:14-35), where every
    trajectory runs its own sampled circuit."""

    def __init__(self, programs, n, device="cuda", dtype=torch.complex64):
        _lib.require_gpu()
        self.device = torch.device(device)
        self.dtype = dtype
        self.code = _lib.dtype_code(dtype)
        self.n = int(n)
        self.programs = list(programs)
        for p in self.programs:
            if p.n != self.n:
                raise ValueError("every program must act on the same number of qubits")
            if any(q.kind != PASS_TILE or q.rb != RB for q in p.passes):
                raise ValueError("batched rounds take tile passes with 4 register bits only")
        blobs, where, off = [], {}, 0
        for p in self.programs:
            if id(p) not in where:
                where[id(p)] = off
                blobs.append((off, p.blob))
                off += (len(p.blob) + 255) & ~255
        host = np.zeros(max(off, 256), dtype=np.uint8)
        for o, b in blobs:
            host[o:o + len(b)] = b
        self.progs = torch.from_numpy(host).to(self.device)
        self.addr = [self.progs.data_ptr() + where[id(p)] for p in self.programs]
        self._tables = None
        self._for_psi = None

    def _plan(self, psi):
        es = 8 if self.dtype == torch.complex64 else 16
        N = 2 ** self.n
        R = max((len(p.passes) for p in self.programs), default=0)
        tables, launches, off = [], [], 0
        for r in range(R):
            groups = {}
            for c, p in enumerate(self.programs):
                if r >= len(p.passes):
                    continue
                h = np.frombuffer(p.blob[:HDR_DT.itemsize].tobytes(), dtype=HDR_DT)[0]
                a = self.addr[c]
                pa = p.passes[r]
                groups.setdefault(pa.tb, []).append((psi.data_ptr() + es * N * c,
                                                     a + int(h['off_passes']) + r * PASS_DT.itemsize,
                                                     a + int(h['off_stages']), a + int(h['off_ops']),
                                                     2 ** (self.n - pa.tb - RB)))
            for tb, items in sorted(groups.items()):
                arr = np.zeros(len(items), dtype=ITEM_DT)
                wg = 0
                for i, (ps, pp, pst, pop, nwg) in enumerate(items):
                    arr[i] = (ps, 0, pp, pst, pop, 0, 0, wg, self.n, 0)
                    wg += nwg
                launches.append((off, len(items), wg, tb))
                tables.append(arr)
                off += (arr.nbytes + 255) & ~255
        blob = np.zeros(max(off, 256), dtype=np.uint8)
        o = 0
        for t in tables:
            blob[o:o + t.nbytes] = np.frombuffer(t.tobytes(), dtype=np.uint8)
            o += (t.nbytes + 255) & ~255
        self._tables = torch.from_numpy(blob).to(self.device)
        self._launches = launches
        self._for_psi = (psi.data_ptr(), psi.numel())

    def run(self, psi):
        """Apply program c to psi[c·2^n : (c+1)·2^n] for every c, in place."""
        if psi.dtype != self.dtype or not psi.is_contiguous() or psi.numel() != len(self.programs) * 2 ** self.n:
            raise ValueError("psi must be a contiguous flat buffer of one state per program")
        if self._for_psi != (psi.data_ptr(), psi.numel()):
            self._plan(psi)
        lib = _lib.load()
        st = _lib.stream_ptr(self.device)
        base = self._tables.data_ptr()
        for off, ni, wg, tb in self._launches:
            _lib.check(lib.svh_batch_tile(base + off, ni, wg, self.code, tb, 0, 0, st), "svh_batch_tile")
