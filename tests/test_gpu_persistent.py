"""The persistent batch (one launch, device task queue) against the oracle.

CPU part (not marked gpu): the host plan itself — every task's dependency is
earlier in its own queue, every unit of every step is covered exactly once,
counters and partial-row buffers are consistent.  GPU part: energies and
gradients of small mixed batches against the oracle's adjoint (c128 ≤ 1e-10
relative, c64 ≤ 1e-5), bitwise-equal graph replays with eager work between
them, a starved grid (2 workgroups: queues drained by other workgroups, long
dependency waits) and circuits whose gates share parameters (slot sums added
in slot order)."""
import numpy as np
import pytest
import torch

from oracle import sv_oracle as O
from qaravan_amd.persistent import PT_ADJ, PT_FWD, PT_HAM, PT_RED, PT_SCT, BatchPlan
from qaravan_amd.program import PASS_DT
from qaravan_amd.workloads import hea_ops, hea_theta, tfim_terms

C64, C128 = torch.complex64, torch.complex128


def _spec(n, layers, params=None):
    spec = []
    k = 0
    for l in range(layers):
        for j, w in enumerate("yz"):
            for q in range(n):
                p = (2 * l + j) * n + q if params is None else params[k]
                spec.append((w, [q], p, None))
                k += 1
        for i in range(n - 1):
            spec.append((None, [i, i + 1], None, np.diag([1, 1, 1, -1]).astype(complex)))
    return spec


def _batch(rng, nc, nmax=13, shared=False):
    ns = [int(x) for x in rng.integers(10, nmax + 1, size=nc)]
    ls = [int(x) for x in rng.integers(1, 5, size=nc)]
    ops, npar, params = [], [], []
    for n, L in zip(ns, ls):
        o = hea_ops(n, L)
        if shared:
            # the RZ layer reuses the RY angles: two gates per parameter
            for op in o:
                if op.kind == 'rot1' and (op.param // n) % 2 == 1:
                    op.param -= n
        params.append([op.param for op in o if op.kind == 'rot1'])
        ops.append(o)
        npar.append(2 * n * L)
    return ns, ls, ops, npar, params


def _check_plan(P):
    pos = {}
    for q in range(P.nq):
        for i in range(int(P.qbeg[q]), int(P.qbeg[q + 1])):
            pos.setdefault(int(P.tasks[i]['desc']), []).append((q, i))
    done_by = {}
    for d, desc in enumerate(P.descs):
        if desc['done'] >= 0:
            assert desc['done'] not in done_by
            done_by[desc['done']] = d
    for d, desc in enumerate(P.descs):
        cover = np.zeros(desc['units'], dtype=int)
        for q, i in pos[d]:
            t = P.tasks[i]
            cover[t['first']:t['first'] + t['count']] += 1
        assert (cover == 1).all()
        if desc['dep'] >= 0 and desc['kind'] != PT_SCT:
            src = done_by[desc['dep']]
            assert desc['dep_need'] == P.descs[src]['units']
            # the waited-for step's tasks all sit earlier in the same queue
            assert {q for q, _ in pos[src]} == {q for q, _ in pos[d]}
            assert max(i for _, i in pos[src]) < min(i for _, i in pos[d])


def _check_bounds(P):
    """Every region a descriptor reads or writes lies inside its buffer (and,
    for states, inside its own circuit's slice)."""
    es = P.es
    size = {'psi': es * int(P.amp_off[-1]), 'lam': es * int(P.amp_off[-1]), 'gpart': 8 * P.gpart_doubles,
            'epart': 8 * int(P.ep_off[-1]), 'terms': len(P.terms_host), 'grad': 8 * int(P.par_off[-1]),
            'energy': 8 * P.nc, 'cidx': 4 * P.nc, 'slotsum': 8 * P.slotsum_doubles, 't_end': 8 * P.nc,
            'trig': es * int(P.par_off[-1]), 'progs': len(P.progs_host)}

    def inside(v, nbytes):
        assert v[0] in size and 0 <= v[1] and v[1] + nbytes <= size[v[0]], (v, nbytes)

    for d in P.descs:
        n = d['n']
        if d['kind'] in (PT_FWD, PT_ADJ, PT_HAM):
            inside(d['psi'], es * 2 ** n)
            c = int(np.searchsorted(P.amp_off, d['psi'][1] // es, side='right')) - 1
            assert d['psi'][1] == es * int(P.amp_off[c]) and P.ns[c] == n
        if d['kind'] in (PT_ADJ, PT_HAM):
            inside(d['lam'], es * 2 ** n)
        if d['kind'] in (PT_FWD, PT_ADJ):
            off = d['pass'][1]
            rec = np.frombuffer(P.progs_host[off:off + PASS_DT.itemsize].tobytes(), dtype=PASS_DT)[0]
            assert d['units'] == 2 ** (n - int(rec['L'])) and d['ng'] == int(rec['n_grad'])
            assert 6 <= int(rec['tb']) <= 9 and int(rec['rb']) == 3
            # λ at its fixed offset, past every ψ tile
            assert 2 ** int(rec['L']) <= P.lam_off
            assert P.max_lds >= es * ((P.lam_off if d['kind'] == PT_ADJ else 0) + 2 ** int(rec['L']))
            if d['kind'] == PT_ADJ:        # one partial row per wave of every tile
                inside(d['part'], 8 * (d['units'] << (int(rec['tb']) - 6)) * max(d['ng'], 1))
        if d['kind'] == PT_HAM:
            inside(d['part'], 8 * d['units'])
            inside(d['out'], 32 * d['ng'])
        if d['kind'] == PT_RED:
            inside(d['part'], 8 * d['rows'] * d['ng'])
            if d.get('slotsum'):
                inside(d['slotsum'], 8 * d['units'])


@pytest.mark.parametrize("threads", [512, 256])
@pytest.mark.parametrize("shared", [False, True])
def test_plan_dependencies_and_coverage(shared, threads):
    rng = np.random.default_rng(7)
    ns, ls, ops, npar, _ = _batch(rng, 10, shared=shared)
    P = BatchPlan(ns, ops, [tfim_terms(n) for n in ns], npar, dtype=C64, nq=3, threads=threads)
    _check_plan(P)
    _check_bounds(P)
    assert P.max_lds >= 8 * threads
    for d in P.descs:            # every tile fits the workgroup
        if d['kind'] in (PT_FWD, PT_ADJ):
            off = d['pass'][1]
            rec = np.frombuffer(P.progs_host[off:off + PASS_DT.itemsize].tobytes(), dtype=PASS_DT)[0]
            assert 2 ** int(rec['tb']) <= threads
    kinds = [d['kind'] for d in P.descs]
    assert kinds.count(PT_HAM) == len(ns) and kinds.count(PT_FWD) > 0 and kinds.count(PT_ADJ) > 0
    assert (kinds.count(PT_SCT) == len(ns)) == shared
    assert P.describe()['tasks'] == len(P.tasks)


def _run(gpu, ns, ops, npar, thetas, dtype, **kw):
    from qaravan_amd.persistent import PersistentBatch
    b = PersistentBatch(ns, ops, [tfim_terms(n) for n in ns], npar, device=gpu, dtype=dtype, **kw)
    b.set_theta(thetas)
    e, g = b.run()
    torch.cuda.synchronize(gpu)
    b.check()
    return b, e.cpu().numpy().copy(), g.cpu().numpy().copy()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,threads", [(C128, 512), (C64, 512), (C64, 256)])
@pytest.mark.parametrize("shared", [False, True])
def test_persistent_vs_oracle(gpu, dtype, threads, shared):
    rng = np.random.default_rng(11 + shared)
    ns, ls, ops, npar, params = _batch(rng, 8, shared=shared)
    thetas = [rng.uniform(-np.pi, np.pi, size=p) for p in npar]
    print(f"persistent batch ns={ns} layers={ls} shared={shared} {dtype} threads={threads}", flush=True)
    b, e, g = _run(gpu, ns, ops, npar, thetas, dtype, nq=3, threads=threads)
    gl = b.gradients()
    tol = 1e-10 if dtype == C128 else 1e-5
    for c, (n, L) in enumerate(zip(ns, ls)):
        e_ref, g_ref = O.adjoint_gradient(_spec(n, L, params[c]), thetas[c], n, O.tfi_terms(n, -1.0, -1.0))
        rel = np.linalg.norm(gl[c] - g_ref) / np.linalg.norm(g_ref)
        assert abs(e[c] - e_ref) < (1e-11 if dtype == C128 else 1e-5), (c, e[c], e_ref)
        assert rel < tol, (c, rel)
    lat = b.latencies_us()
    assert (lat > 0).all()


@pytest.mark.gpu
def test_persistent_graph_replays_and_starved_grid(gpu):
    rng = np.random.default_rng(5)
    ns, ls, ops, npar, params = _batch(rng, 12, nmax=14)
    thetas = [rng.uniform(-np.pi, np.pi, size=p) for p in npar]
    print(f"persistent replays ns={ns}", flush=True)
    b, e0, g0 = _run(gpu, ns, ops, npar, thetas, C64)
    b.capture()
    for r in range(3):
        e, g = b.run()
        torch.cuda.synchronize(gpu)
        b.check()
        assert np.array_equal(e.cpu().numpy(), e0) and np.array_equal(g.cpu().numpy(), g0), r
        x = torch.randn(1 << 20, device=gpu)       # eager work between replays
        (x * 2).sum().item()
    # two workgroups for the whole batch: every queue but two is served by a
    # workgroup whose own queue is drained
    b2, e2, g2 = _run(gpu, ns, ops, npar, thetas, C64, max_workgroups=2, plan=b.plan)
    assert np.array_equal(e2, e0) and np.array_equal(g2, g0)
