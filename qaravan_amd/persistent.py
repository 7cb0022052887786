"""Persistent batched forward + adjoint: a whole batch of parametrised circuits
in ONE kernel launch with a device task queue.

qaravan evaluates many circuits one simulation at a time (the trajectories of
src/qaravan/tensorQ/monte_carlo_sim.py:14-35, one StatevectorSim per input
state and step in applications/compilation.py:73-133).  Here every circuit c of
a batch, with its own θ_c and Pauli-sum loss E_c = ⟨ψ_c|H_c|ψ_c⟩, is a short
dependency chain of steps:

    forward pass 0 → … → forward pass F-1 → λ = Hψ (blocks) → adjoint pass 0 → …
                                              ↘ energy sum          ↘ gradient sums of each pass

Each step is a descriptor (svh_pdesc_t) whose units (tiles, Hamiltonian
blocks, reduction slots) are cut into tasks.  ``svh_batch_persistent`` runs a
resident grid of workgroups (512 threads, or 256 for complex64) that take
tasks from per-queue lists
and wait, per task, only for the one step it reads (a device counter), so a
circuit advances as soon as its own previous step is done: small circuits
finish early instead of waiting for the slowest pass of a round, and no launch
gaps separate the passes.

Queues: circuits are spread over ``nq`` queues (one per XCD by default:
workgroup b serves queue b mod nq, and workgroups are placed round-robin on
the XCDs, so a circuit's state stays in one XCD's L2 while it is worked on),
balanced by estimated cost.  Inside a queue the task order comes from a list-
scheduling simulation of the queue's workers: among the tasks whose step is
ready, the cheapest circuit first (shortest job first, the order that
minimises mean completion time); every task's dependency is therefore earlier
in its own queue, which is what makes the device waits finite.

The tile passes run in the generic kernels (qaravan_amd.patterns), so the
batch needs no compiler at run time.  Everything is planned on the host
without a GPU (``BatchPlan``); ``PersistentBatch`` binds a plan to device
buffers, and ``run()`` is one library call, recorded in a HIP graph by
``capture()``.
"""
import heapq

import numpy as np
import torch

from . import _lib
from .adjoint import ParamProgram, term_table
from .batch import batch_config
from .program import HDR_DT, OPF_GRAD, PASS_DT, PASS_TILE, STAGE_DT

PT_FWD, PT_ADJ, PT_HAM, PT_RED, PT_SCT = 1, 2, 3, 4, 5
PDESC_DT = np.dtype([('kind', '<i4'), ('n', '<i4'), ('dep', '<i4'), ('dep_need', '<i4'), ('done', '<i4'),
                     ('units', '<i4'), ('ng', '<i4'), ('rows', '<i4'), ('fin', '<i4'), ('fin_need', '<i4'),
                     ('psi', '<u8'), ('lam', '<u8'), ('pass', '<u8'), ('stages', '<u8'), ('ops', '<u8'),
                     ('trig', '<u8'), ('part', '<u8'), ('out', '<u8'), ('slot_param', '<u8'),
                     ('slotsum', '<u8'), ('tstamp', '<u8')])
PTASK_DT = np.dtype([('desc', '<i4'), ('first', '<i4'), ('count', '<i4'), ('pad', '<i4')])
assert PDESC_DT.itemsize == 128 and PTASK_DT.itemsize == 16
_PTR_FIELDS = ('psi', 'lam', 'pass', 'stages', 'ops', 'trig', 'part', 'out', 'slot_param', 'slotsum', 'tstamp')
CUS_PER_XCD = 32         # MI355X: 256 CUs in 8 XCDs


def queue_workers(threads):
    """Resident workgroups per XCD (the schedule simulation's worker count):
    2 per CU of 512 threads, 4 of 256."""
    return CUS_PER_XCD * (1024 // threads)


def _hdr(blob):
    return np.frombuffer(blob[:HDR_DT.itemsize].tobytes(), dtype=HDR_DT)[0]


class BatchPlan:
    """Host plan of a persistent batch; needs no GPU.

    Pointer fields of the descriptors are (buffer name, byte offset) pairs
    until ``bind`` turns them into addresses.  Buffers: psi, lam (states,
    amplitude offsets ``amp_off``), progs (every distinct program blob),
    trig / theta / grad (parameter offsets ``par_off``), gpart (gradient
    partial rows), epart (energy partials), terms, energy, cidx, slotsum,
    t_end."""

    def __init__(self, ns, op_lists, terms_list, n_params_list, dtype=torch.complex64, keys=None, nq=8,
                 remap_adjoint=None, workers_per_queue=None, tasks_per_step=None, threads=None):
        self.dtype = dtype
        from .autotune import batch_tuning
        tuned = batch_tuning(dtype)
        # workgroup size (the tuned value, else 512; 256 only for complex64):
        # every tile pass has at most this many threads
        self.threads = int(threads if threads is not None else tuned.get("threads", 512))
        if self.threads not in (256, 512) or (self.threads == 256 and dtype != torch.complex64):
            raise ValueError("persistent batch workgroups: 512 threads, or 256 for complex64")
        # at most this many tasks per step (the tuned value, else 128)
        self.tasks_per_step = int(tasks_per_step if tasks_per_step is not None else tuned.get("tasks_per_step", 128))

        self.es = 8 if dtype == torch.complex64 else 16
        self.ns = [int(n) for n in ns]
        self.nc = len(self.ns)
        if min(self.ns) < 10:
            raise NotImplementedError("the persistent batch works on Hamiltonian blocks of 1024 amplitudes: n >= 10")
        self.nq = max(1, min(int(nq), self.nc))
        # ---- programs (shared by circuits with equal keys)
        self.pps, cache = [], {}
        for c in range(self.nc):
            key = None if keys is None else keys[c]
            if key is not None and key in cache:
                self.pps.append(cache[key])
                continue
            # remapped adjoint sweeps: as asked (a bool, or {str(n): bool}), else
            # the tuned choice for this n, else on
            per_n = remap_adjoint if isinstance(remap_adjoint, dict) else tuned.get("remap_adjoint", {})
            rm = remap_adjoint if isinstance(remap_adjoint, bool) else per_n.get(str(self.ns[c]), True)
            cfg, adj_cfg = batch_config(self.ns[c], dtype, 3, rm, self.threads)
            pp = ParamProgram(None, terms_list[c], n_params_list[c], cfg=cfg, adj_cfg=adj_cfg, dtype=dtype,
                              ops=op_lists[c], n=self.ns[c])
            if pp.layout is not None or pp.fwd.final_pos is not None:
                raise ValueError("batched forward programs keep the natural bit order")
            if key is not None:
                cache[key] = pp
            self.pps.append(pp)
        # ---- buffer layout
        self.amp_off = np.concatenate([[0], np.cumsum([2 ** n for n in self.ns])]).astype(np.int64)
        self.par_off = np.concatenate([[0], np.cumsum(n_params_list)]).astype(np.int64)
        self.prog_off, blobs, off = {}, [], 0
        for pp in self.pps:
            for prog in (pp.fwd, pp.adj):
                if id(prog) not in self.prog_off:
                    self.prog_off[id(prog)] = off
                    blobs.append((off, prog.blob))
                    off += (len(prog.blob) + 255) & ~255
        self.progs_host = np.zeros(max(off, 256), dtype=np.uint8)
        for o, b in blobs:
            self.progs_host[o:o + len(b)] = b
        tt = [term_table(self.ns[c], terms_list[c]) for c in range(self.nc)]
        self.t_off = np.concatenate([[0], np.cumsum([len(t) for t in tt])]).astype(np.int64)
        self.terms_host = np.concatenate(tt).view(np.uint8).copy()
        self.ep_off = np.concatenate([[0], np.cumsum([2 ** n // 1024 for n in self.ns])]).astype(np.int64)
        self._build_descs()
        self._build_tasks(workers_per_queue or queue_workers(self.threads))

    # ------------------------------------------------------------ descriptors
    def _build_descs(self):
        descs, self.chains = [], []
        self.gpart_doubles = 0
        self.slotsum_doubles = 0
        self.ncounters = 0
        self.max_lds = self.threads * 8      # the reductions' tree (one double per thread)
        self.lam_off = self.threads * (8 if self.dtype == torch.complex64 else 4)
        self.cost = np.zeros(self.nc)

        def counter():
            self.ncounters += 1
            return self.ncounters - 1

        def add(**kw):
            d = dict(kind=0, n=0, dep=-1, dep_need=0, done=-1, units=0, ng=0, rows=0, fin=-1, fin_need=0)
            d.update(kw)
            descs.append(d)
            return len(descs) - 1

        for c in range(self.nc):
            pp, n = self.pps[c], self.ns[c]
            psi, lam = ('psi', self.es * int(self.amp_off[c])), ('lam', self.es * int(self.amp_off[c]))
            trig = ('trig', self.es * int(self.par_off[c]))
            chain = []            # (desc, [dependency descs]) in a valid order
            prev, prev_need, prev_desc = -1, 0, None

            def tile_desc(prog, pi, dual, dep, need, dep_desc, part=None):
                base = self.prog_off[id(prog)]
                h = _hdr(prog.blob)
                p = prog.passes[pi]
                L = p.tb + p.rb
                assert 2 ** p.tb <= self.threads, "a tile pass wider than the workgroup"
                nwg = 2 ** (n - L)
                ps = _pass_row(prog, pi)
                ng = int(ps['n_grad'])
                done = counter()
                d = add(kind=PT_ADJ if dual else PT_FWD, n=n, dep=dep, dep_need=need, done=done, units=nwg, ng=ng,
                        psi=psi, lam=lam if dual else 0,
                        **{'pass': ('progs', base + int(h['off_passes']) + pi * PASS_DT.itemsize)},
                        stages=('progs', base + int(h['off_stages']) + int(ps['stage_begin']) * STAGE_DT.itemsize),
                        ops=('progs', base + int(h['off_ops'])), trig=trig, part=part or 0)
                # λ sits at a fixed offset (threads·8 complex64 / threads·4 complex128 amplitudes)
                self.max_lds = max(self.max_lds, self.es * ((self.lam_off if dual else 0) + 2 ** L))
                unit = (1.0 + 0.12 * len(p.stages) + 0.04 * len(p.ops) * (2.5 if dual else 1.0)) * 2 ** (L - 12)
                self.cost[c] += unit * nwg
                chain.append((d, [dep_desc] if dep_desc is not None else [], unit))
                return d, done, nwg

            for pi in range(len(pp.fwd.passes)):
                prev_desc, prev, prev_need = tile_desc(pp.fwd, pi, False, prev, prev_need, prev_desc)
            # λ = Hψ and the energy partials
            nblk = 2 ** n // 1024
            ham_done = counter()
            ham = add(kind=PT_HAM, n=n, dep=prev, dep_need=prev_need, done=ham_done, units=nblk,
                      ng=int(self.t_off[c + 1] - self.t_off[c]), psi=psi, lam=lam,
                      part=('epart', 8 * int(self.ep_off[c])), out=('terms', 32 * int(self.t_off[c])))
            hunit = 0.15 + 0.01 * int(self.t_off[c + 1] - self.t_off[c])
            chain.append((ham, [prev_desc], hunit))
            self.cost[c] += hunit * nblk
            fin = counter()
            # adjoint passes and their gradient sums
            adj = pp.adj
            slots = [int(x) for x in _grad_slots(adj)]
            distinct = len(set(slots)) == len(slots)
            fin_need = 1 + len(slots)
            sdoff = self.slotsum_doubles
            if not distinct:
                self.slotsum_doubles += len(slots)
            t_end = ('t_end', 8 * c)
            ered = add(kind=PT_RED, n=n, dep=ham_done, dep_need=nblk, units=1, ng=1, rows=nblk,
                       fin=fin, fin_need=fin_need, part=('epart', 8 * int(self.ep_off[c])),
                       out=('energy', 0), slot_param=('cidx', 4 * c), tstamp=t_end if distinct else 0)
            chain.append((ered, [ham], 0.05 + 0.002 * nblk))
            prev, prev_need, prev_desc = ham_done, nblk, ham
            reds = [ered]
            h = _hdr(adj.blob)
            for pi in range(len(adj.passes)):
                ps = _pass_row(adj, pi)
                ng = int(ps['n_grad'])
                p = adj.passes[pi]
                rows = 2 ** (n - p.tb - p.rb) << (p.tb - 6)
                part = ('gpart', 8 * self.gpart_doubles)
                self.gpart_doubles += rows * max(ng, 1)
                d, done, nwg = tile_desc(adj, pi, True, prev, prev_need, prev_desc, part)
                if ng:
                    gb = int(ps['grad_begin'])
                    red = add(kind=PT_RED, n=n, dep=done, dep_need=nwg, units=ng, ng=ng, rows=rows,
                              fin=fin, fin_need=fin_need, part=part, out=('grad', 8 * int(self.par_off[c])),
                              slot_param=('progs', self.prog_off[id(adj)] + int(h['off_grad']) + 4 * gb),
                              slotsum=('slotsum', 8 * (sdoff + gb)) if not distinct else 0,
                              tstamp=t_end if distinct else 0)
                    chain.append((red, [d], 0.05 + 0.0005 * rows))
                    reds.append(red)
                prev, prev_need, prev_desc = done, nwg, d
            if not distinct:
                sct = add(kind=PT_SCT, n=n, dep=fin, dep_need=fin_need, units=1, ng=len(slots),
                          out=('grad', 8 * int(self.par_off[c])),
                          slot_param=('progs', self.prog_off[id(adj)] + int(h['off_grad'])),
                          slotsum=('slotsum', 8 * sdoff), tstamp=t_end)
                chain.append((sct, reds, 0.05))
            self.chains.append(chain)
        self.descs = descs

    # ------------------------------------------------------------ tasks
    def _build_tasks(self, W):
        # circuits to queues: longest processing time first onto the least loaded queue
        load = [(0.0, q) for q in range(self.nq)]
        heapq.heapify(load)
        self.queue_of = np.zeros(self.nc, dtype=np.int32)
        for c in np.argsort(-self.cost, kind="stable"):
            l, q = heapq.heappop(load)
            self.queue_of[c] = q
            heapq.heappush(load, (l + self.cost[c], q))
        queues = [self._schedule([c for c in range(self.nc) if self.queue_of[c] == q], W) for q in range(self.nq)]
        self.qbeg = np.concatenate([[0], np.cumsum([len(t) for t in queues])]).astype(np.int32)
        self.tasks = np.zeros(int(self.qbeg[-1]), dtype=PTASK_DT)
        for q, tl in enumerate(queues):
            for i, (d, first, count) in enumerate(tl):
                self.tasks[int(self.qbeg[q]) + i] = (d, first, count, 0)

    def _schedule(self, circs, W):
        """Task order of one queue: list scheduling of W workers, shortest
        circuit first among the ready tasks.  Returns [(desc, first, count)]."""
        if not circs:
            return []
        prio = {c: r for r, c in enumerate(sorted(circs, key=lambda c: self.cost[c]))}
        order = []
        finish = {}          # desc -> completion time (known once all its tasks are placed)
        pending = {}         # desc -> [tasks left, latest end]
        waiting = {}         # desc -> dependencies not yet complete
        future = []          # heap of (ready time, prio, chain position, desc, first, count, unit cost)
        avail = []           # heap of (prio, chain position, desc, first, count, unit cost, ready time)
        users, deps_of, info = {}, {}, {}

        def chunk(units):
            # at most tasks_per_step tasks per step: enough to spread a large
            # pass over the workers (a 22-qubit pass of 1024 tiles feeds all 64
            # workers of its queue), few enough that queue claims, fences and
            # counter updates stay rare (128: 575 -> 515 ms for the 512-circuit
            # batch against 32, with the workers' dependency waits 13 % -> 3 %
            # of their time)
            return max(1, -(-units // self.tasks_per_step))

        for c in circs:
            for pos, (d, deps, unit) in enumerate(self.chains[c]):
                info[d] = (c, pos, unit)
                waiting[d] = len(deps)
                deps_of[d] = deps
                for dd in deps:
                    users.setdefault(dd, []).append(d)

        def release(d, t):
            c, pos, unit = info[d]
            units = self.descs[d]['units']
            k = chunk(units)
            cnt = 0
            for first in range(0, units, k):
                heapq.heappush(future, (t, prio[c], pos, d, first, min(k, units - first), unit))
                cnt += 1
            pending[d] = [cnt, t]

        for c in circs:
            release(self.chains[c][0][0], 0.0)
        workers = [0.0] * W
        while future or avail:
            t_free = heapq.heappop(workers)
            if not avail and future[0][0] > t_free:
                t_free = future[0][0]           # the worker idles until something is ready
            while future and future[0][0] <= t_free:
                t, pr, pos, d, first, count, unit = heapq.heappop(future)
                heapq.heappush(avail, (pr, pos, d, first, count, unit, t))
            _, _, d, first, count, unit, t_ready = heapq.heappop(avail)
            end = max(t_free, t_ready) + unit * count
            heapq.heappush(workers, end)
            order.append((d, first, count))
            pd = pending[d]
            pd[0] -= 1
            pd[1] = max(pd[1], end)
            if pd[0] == 0:
                finish[d] = pd[1]
                for u in users.get(d, []):
                    waiting[u] -= 1
                    if waiting[u] == 0:
                        release(u, max(finish[x] for x in deps_of[u]))
        return order

    # ------------------------------------------------------------ binding
    def bind(self, base):
        """Descriptor table with addresses, for buffer base addresses ``base``."""
        arr = np.zeros(len(self.descs), dtype=PDESC_DT)
        for i, d in enumerate(self.descs):
            row = arr[i]
            for k, v in d.items():
                if k in _PTR_FIELDS:
                    row[k] = 0 if v == 0 or v is None else base[v[0]] + int(v[1])
                else:
                    row[k] = v
        return arr

    def describe(self):
        kinds = np.array([d['kind'] for d in self.descs])
        return dict(circuits=self.nc, queues=self.nq, descriptors=len(self.descs), tasks=len(self.tasks),
                    tile_units=int(sum(d['units'] for d in self.descs if d['kind'] in (PT_FWD, PT_ADJ))),
                    fwd_passes=int((kinds == PT_FWD).sum()), adj_passes=int((kinds == PT_ADJ).sum()),
                    counters=self.ncounters, lds_bytes=self.max_lds)

    def algorithmic_bytes(self):
        """Bytes one run must move at least: every tile pass reads and writes
        its state (ψ, and λ in the adjoint sweep) once, the Hamiltonian step
        reads ψ once and writes λ."""
        b = 0
        for d in self.descs:
            amps = 2 ** d['n']
            if d['kind'] == PT_FWD:
                b += 2 * amps * self.es
            elif d['kind'] == PT_ADJ:
                b += 4 * amps * self.es
            elif d['kind'] == PT_HAM:
                b += 2 * amps * self.es
        return b


def _pass_row(prog, pi):
    h = _hdr(prog.blob)
    off = int(h['off_passes']) + pi * PASS_DT.itemsize
    return np.frombuffer(prog.blob[off:off + PASS_DT.itemsize].tobytes(), dtype=PASS_DT)[0]


def _grad_slots(prog):
    h = _hdr(prog.blob)
    off = int(h['off_grad'])
    return np.frombuffer(prog.blob[off:off + 4 * int(h['n_grad_slots'])].tobytes(), dtype=np.int32)


class PersistentBatch:
    """Energies and gradients of a batch of parametrised circuits, computed by
    one persistent launch (see the module docstring).  Same inputs as
    ``BatchPlan``; ``set_theta`` then ``run()`` → (energies, flat gradient)."""

    def __init__(self, ns, op_lists, terms_list, n_params_list, device="cuda", dtype=torch.complex64, keys=None,
                 nq=8, remap_adjoint=None, max_workgroups=0, plan=None, xcd_local=True, threads=None):
        _lib.require_gpu()
        self.device = torch.device(device)
        self.dtype = dtype
        self.code = _lib.dtype_code(dtype)
        self.plan = plan or BatchPlan(ns, op_lists, terms_list, n_params_list, dtype, keys, nq, remap_adjoint,
                                      threads=threads)
        P = self.plan
        dev = self.device
        self.nc = P.nc
        self.max_workgroups = int(max_workgroups)
        self.xcd_local = bool(xcd_local)
        self.sync_mode = None       # set by the first launch: 1 = XCD-local, 0 = device-wide
        total = int(P.amp_off[-1])
        npar = int(P.par_off[-1])
        self.psi = torch.empty(total, dtype=dtype, device=dev)
        self.lam = torch.empty(total, dtype=dtype, device=dev)
        self.theta = torch.zeros(npar, dtype=torch.float64, device=dev)
        self.trig = torch.empty(npar, dtype=dtype, device=dev)
        self.grad = torch.zeros(npar, dtype=torch.float64, device=dev)
        self.energy = torch.zeros(self.nc, dtype=torch.float64, device=dev)
        self.gpart = torch.zeros(max(P.gpart_doubles, 1), dtype=torch.float64, device=dev)
        self.epart = torch.zeros(max(int(P.ep_off[-1]), 1), dtype=torch.float64, device=dev)
        self.slotsum = torch.zeros(max(P.slotsum_doubles, 1), dtype=torch.float64, device=dev)
        self.progs = torch.from_numpy(P.progs_host).to(dev)
        self.terms = torch.from_numpy(P.terms_host).to(dev)
        self.cidx = torch.arange(self.nc, dtype=torch.int32, device=dev)
        self.t0 = torch.zeros(1, dtype=torch.int64, device=dev)
        self.t_end = torch.zeros(self.nc, dtype=torch.int64, device=dev)
        self.offsets = torch.tensor(P.amp_off[:-1], dtype=torch.int64, device=dev)
        base = {k: getattr(self, k).data_ptr() for k in ('psi', 'lam', 'trig', 'grad', 'energy', 'gpart', 'epart',
                                                         'slotsum', 'progs', 'terms', 'cidx', 't_end')}
        self.descs = torch.from_numpy(P.bind(base).view(np.uint8).copy()).to(dev)
        self.tasks = torch.from_numpy(P.tasks.view(np.uint8).copy()).to(dev)
        self.qbeg = torch.from_numpy(P.qbeg).to(dev)
        self.heads = torch.zeros(P.nq, dtype=torch.int32, device=dev)
        self.counters = torch.zeros(max(P.ncounters, 1), dtype=torch.int32, device=dev)
        self.err = torch.zeros(1, dtype=torch.int32, device=dev)
        self.stats = None           # enable_stats(): scheduler time accounting
        self.graph = None

    def enable_stats(self):
        """Have every launch add the workers' wait and task times (see
        ``scheduler_stats``); call before ``capture()``."""
        self.stats = torch.zeros(5, dtype=torch.int64, device=self.device)

    def scheduler_stats(self):
        """Accumulated since enable_stats(): seconds the workers spent waiting
        on dependencies and running tasks (waits included), task count, and
        waits that had to spin, and the workgroups of the grid (summed over
        the launches)."""
        w, t, n, s, g = (int(x) for x in self.stats.cpu().tolist())
        return dict(wait_s=w / 1e8, task_s=t / 1e8, tasks=n, waits_spun=s, workgroups=g)

    def _launch(self):
        lib = _lib.load()
        P = self.plan
        st = _lib.stream_ptr(self.device)
        self.grad.zero_()
        self.energy.zero_()
        self.t_end.zero_()
        _lib.check(lib.svh_timestamp(_lib.ptr(self.t0), st), "svh_timestamp")
        _lib.check(lib.svh_batch_init(_lib.ptr(self.psi), self.psi.numel(), _lib.ptr(self.offsets), self.nc,
                                      self.code, st), "svh_batch_init")
        _lib.check(lib.svh_batch_trig(_lib.ptr(self.theta), self.theta.numel(), _lib.ptr(self.trig), self.code, st),
                   "svh_batch_trig")
        import ctypes
        mode = ctypes.c_int(-1)
        _lib.check(lib.svh_batch_persistent(_lib.ptr(self.descs), _lib.ptr(self.tasks), _lib.ptr(self.qbeg), P.nq,
                                            _lib.ptr(self.heads), _lib.ptr(self.counters), P.ncounters,
                                            _lib.ptr(self.err), self.code, P.max_lds, P.threads,
                                            self.max_workgroups,
                                            int(self.xcd_local), ctypes.byref(mode),
                                            _lib.ptr(self.stats) if self.stats is not None else None, st),
                   "svh_batch_persistent")
        self.sync_mode = mode.value

    def set_theta(self, thetas):
        """θ for every circuit (list of arrays/tensors, or one flat tensor)."""
        if isinstance(thetas, torch.Tensor) and thetas.dim() == 1 and thetas.numel() == self.theta.numel():
            self.theta.copy_(thetas.to(torch.float64))
        else:
            flat = np.concatenate([np.asarray(t.detach().cpu() if hasattr(t, "detach") else t, dtype=np.float64)
                                   for t in thetas])
            self.theta.copy_(torch.from_numpy(flat))

    def run(self):
        """One batch (graph replay when captured); returns (energies, flat gradient)."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch()
        return self.energy, self.grad

    def check(self):
        """Raise if a device dependency wait ran past its bound, or a queue was
        not fully taken (synchronises)."""
        if int(self.err.item()) != 0:
            what = ("a dependency wait exceeded its bound" if int(self.err.item()) == 1
                    else "a tile pass is wider than the workgroup")
            raise _lib.HipLibraryError(f"persistent batch: {what}; results invalid")
        heads = self.heads.cpu().numpy()
        qlen = np.diff(self.plan.qbeg)
        if (heads < qlen).any():
            raise _lib.HipLibraryError(f"persistent batch: queues not drained ({heads.tolist()} of "
                                       f"{qlen.tolist()}); results invalid")

    def capture(self):
        self._launch()
        torch.cuda.synchronize(self.device)
        self.check()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._launch()
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def gradients(self):
        g = self.grad.cpu().numpy()
        P = self.plan
        return [g[P.par_off[c]:P.par_off[c + 1]] for c in range(self.nc)]

    def latencies_us(self):
        """Per-circuit completion time since the batch started (device clock, 100 MHz)."""
        t0 = int(self.t0.item())
        return (self.t_end.cpu().numpy().astype(np.int64) - t0) / 100.0
