"""Program specialisation: one generated kernel per tile pass.

The generic tile kernel interprets its op list: every op is a run-time switch
over (kind, register bits).  On gfx950 that costs about a register-file copy
per op at the switch merge.  A program is fixed once compiled (θ and matrices
stay in device memory), so its forward passes can instead be spelled out as
straight-line code: this module writes one ``svh_sp_<p>`` kernel per tile
pass, built from the device pieces in csrc/svh_tile_ops.h, compiles the file
with hipcc into a code object (cached by content hash), and the library
launches those kernels for the passes they cover (``svh_run_program_sp``).
"""
import hashlib
import os
import sys
import time
import subprocess

import numpy as np

from . import program as P

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
# compiled objects are written to CACHE_DIR and looked up there and in the
# package's own cache (objects built ahead of time, e.g. by build())
PACKAGE_CACHE = os.path.join(HERE, "_sp_cache")
CACHE_DIR = os.environ.get("QARAVAN_SP_CACHE", PACKAGE_CACHE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QARAVAN_HIP_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "--genco",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + CSRC]
MAX_LDS = 160 * 1024          # LDS per workgroup on gfx950

# cold-start accounting: code objects compiled by this process and the wall
# time of the compile calls that had to run hipcc
STATS = {"objects_compiled": 0, "compile_wall_s": 0.0}


def _decode(blob):
    h = np.frombuffer(blob[:P.HDR_DT.itemsize].tobytes(), dtype=P.HDR_DT)[0]

    def arr(name, dt, count):
        off = int(h['off_' + name])
        return np.frombuffer(blob[off:off + dt.itemsize * count].tobytes(), dtype=dt)
    return (h, arr('passes', P.PASS_DT, int(h['n_passes'])), arr('stages', P.STAGE_DT, int(h['n_stages'])),
            arr('ops', P.OP_DT, int(h['n_ops'])))


def lds_bytes(p, dtype_code, dual):
    es = 8 if dtype_code == 0 else 16
    nw = max(1, (2 ** int(p['tb'])) // 64)
    amps = 2 ** (int(p['L']) - lds_split(p, dual))
    return (2 if dual else 1) * amps * es + (8 * nw * int(p['n_grad']) if dual else 0)


def lds_split(p, dual=False):
    """log2 of the parts a pass's LDS re-distributions go in."""
    f = int(p['flags'])
    return 1 if f & P.PASSF_HALF_LDS else 2 if f & P.PASSF_QUARTER_LDS else 0


def _lds_swz(x):
    return x ^ ((x >> 5) & 31)


def split_bits(P_rbits, S_rbits, rb, k):
    """The k highest tile-local bits both layouts hold in registers, highest first."""
    common = set(int(x) for x in P_rbits[:rb]) & set(int(x) for x in S_rbits[:rb])
    return sorted(common, reverse=True)[:k]


def drop_bits(x, hs):
    """x with the bits hs (highest first) removed, the bits above shifted down."""
    for h in hs:
        x = (x & ((1 << h) - 1)) | ((x >> (h + 1)) << h)
    return x


def _split_exchange(R, rb, P_rbits, S_rbits, first, sp, sc, k, dual=False):
    """One LDS re-distribution from stage sp's layout (register bits P_rbits)
    to stage sc's (S_rbits) in 2^k parts (dual: ψ in v through lds and λ in
    w through lds2, slot for slot).  The parts split on hs, the k
    highest tile-local bits that both layouts hold in registers: part q moves
    the elements whose bits hs read q, each part through the same 2^(L-k) LDS
    slots at swizzle(index with the bits hs removed).  Removal and swizzle are
    linear over GF(2), so a slot is (thread part) ^ (constant per register);
    the constants are computed here."""
    P_rbits, S_rbits = [int(x) for x in P_rbits[:rb]], [int(x) for x in S_rbits[:rb]]
    hs = split_bits(P_rbits, S_rbits, rb, k)
    jp = [P_rbits.index(h) for h in hs]
    js = [S_rbits.index(h) for h in hs]

    def roff(rbits, j, skip):
        x = 0
        for kk in range(rb):
            if kk not in skip and (j >> kk) & 1:
                x |= 1 << rbits[kk]
        return _lds_swz(drop_bits(x, hs))

    def part(j, idx):
        return sum(((j >> b) & 1) << i for i, b in enumerate(idx))

    def tpart(s):
        x = f"thr_local<{rb}>(st[{s}], threadIdx.x)"
        for h in hs:
            x = f"(({x}) & {(1 << h) - 1}) | ((({x}) >> {h + 1}) << {h})"
        return f"lds_swz({x})"

    arrays = [("v", "lds", "x"), ("w", "lds2", "y")] if dual else [("v", "lds", "x")]
    out = ["    {",
           f"        const int tp = {tpart(sp)};",
           f"        const int tn = {tpart(sc)};"]
    out += [f"        cpx<{R}> {t}[{1 << rb}];" for _, _, t in arrays]
    for q in range(1 << len(hs)):
        if not (first and q == 0):
            out.append("        __syncthreads();")
        for reg, buf, _ in arrays:
            out += [f"        {buf}[tp ^ {roff(P_rbits, j, jp)}] = {reg}[{j}];"
                    for j in range(1 << rb) if part(j, jp) == q]
        out.append("        __syncthreads();")
        for _, buf, t in arrays:
            out += [f"        {t}[{j}] = {buf}[tn ^ {roff(S_rbits, j, js)}];"
                    for j in range(1 << rb) if part(j, js) == q]
    for reg, _, t in arrays:
        out.append(f"        for (int j = 0; j < {1 << rb}; ++j) {reg}[j] = {t}[j];")
    out.append("    }")
    return out


def eligible(p, dtype_code, dual=False):
    """Forward kernels take 3 to 5 register bits; adjoint kernels 3 or 4."""
    rb = int(p['rb'])
    return (int(p['kind']) == P.PASS_TILE and (rb in (3, 4) or (rb == 5 and not dual))
            and lds_bytes(p, dtype_code, dual) <= MAX_LDS and (dual or int(p['n_grad']) == 0))


def _body(R, p, stages, ops, dual, store_to="psi"):
    """Statements of one pass after the prologue has defined psi (and lam),
    n, pass, st (this pass's stages), ops (this pass's ops), trig, gpart and
    wg (the tile index)."""
    tb, rb = int(p['tb']), int(p['rb'])
    s0, ns = int(p['stage_begin']), int(p['n_stages'])
    op0 = int(stages[s0]['op_begin'])
    ng = int(p['n_grad'])
    L = int(p['L'])
    out = ["    extern __shared__ unsigned char smem_raw[];",
           f"    cpx<{R}> *lds = reinterpret_cast<cpx<{R}> *>(smem_raw);"]
    if dual:
        out += [f"    cpx<{R}> *lds2 = lds + {2 ** (L - lds_split(p, dual))};",
                f"    double *gacc = reinterpret_cast<double *>(lds2 + {2 ** (L - lds_split(p, dual))});",
                f"    cpx<{R}> v[{1 << rb}], w[{1 << rb}];",
                f"    const uint64_t base = sp_base_wg<{R}, {tb}, {rb}>(pass, n, wg);",
                f"    sp_load2<{R}, {tb}, {rb}>(psi, lam, base, st[0], v, w);"]
    else:
        # the first pass of a program that starts from |0…0⟩ generates it
        load = "sp_gen_zero" if int(p['flags']) & P.PASSF_GEN_ZERO else "sp_load"
        src = "" if load == "sp_gen_zero" else "psi, "
        out += [f"    cpx<{R}> v[{1 << rb}];",
                f"    const uint64_t base = sp_base_wg<{R}, {tb}, {rb}>(pass, n, wg);",
                f"    {load}<{R}, {tb}, {rb}>({src}base, st[0], v);"]
    W = min(1 << tb, 64)
    grads = []          # (variable, op index) of this stage's gradient terms not yet stored

    def flush():
        # the stage's gradient terms (consecutive slots) summed over the wave
        # together, up to 8 at a time
        if not grads:
            return
        cnt = len(grads)
        slots = [int(ops[op0 + k]['grad_slot']) for _, k in grads]
        if slots != list(range(slots[0], slots[0] + cnt)):
            raise ValueError(f"gradient slots of a stage are not consecutive: {slots}")
        K = (cnt - 1).bit_length()
        vals = ", ".join([g for g, _ in grads] + ["0.0"] * ((1 << K) - cnt))
        out.append(f"    {{ double gv[{1 << K}] = {{{vals}}}; "
                   f"sp_grad_group<{K}, {W}>(gv, gacc, {ng}, ops[{grads[0][1]}].grad_slot, {cnt}); }}")
        grads.clear()

    for s in range(ns):
        S = stages[s0 + s]
        if s > 0:
            first = "true" if s == 1 else "false"
            if lds_split(p, dual):
                out += _split_exchange(R, rb, stages[s0 + s - 1]['rbits'], S['rbits'], s == 1, s - 1, s,
                                       lds_split(p, dual), dual)
            else:
                out.append(f"    sp_exchange2<{R}, {first}, {rb}>(lds, lds2, st[{s - 1}], st[{s}], v, w);" if dual
                           else f"    sp_exchange<{R}, {first}, {rb}>(lds, st[{s - 1}], st[{s}], v);")
        for oi in range(int(S['op_begin']), int(S['op_end'])):
            o = ops[oi]
            k = oi - op0
            kind, a, b, pauli = int(o['kind']), int(o['a']), int(o['b']), int(o['pauli'])
            hi, lo = max(a, b), min(a, b)
            if dual:
                g = "true" if int(o['flags']) & P.OPF_GRAD else "false"
                if kind == P.OP_MAT1:
                    out.append(f"    sp_mat1<{R}, {a}, {rb}>(v, ops[{k}]); sp_mat1<{R}, {a}, {rb}>(w, ops[{k}]);")
                elif kind == P.OP_MAT2:
                    out.append(f"    sp_mat2<{R}, {a}, {b}, {rb}>(v, ops[{k}]); "
                               f"sp_mat2<{R}, {a}, {b}, {rb}>(w, ops[{k}]);")
                elif kind == P.OP_CX:
                    th = "true" if a > b else "false"
                    out.append(f"    ap_cx_t<{R}, {rb}, {hi}, {lo}, {th}>(v); ap_cx_t<{R}, {rb}, {hi}, {lo}, {th}>(w);")
                elif kind == P.OP_CZ:
                    out.append(f"    ap_cz<{R}, {rb}, {hi}, {lo}>(v); ap_cz<{R}, {rb}, {hi}, {lo}>(w);")
                elif kind in (P.OP_ROT1, P.OP_ROT2):
                    if g == "true":
                        var = f"gq{oi - op0}"
                        gr = (f"gr_rot1<{R}, {rb}, {a}>" if kind == P.OP_ROT1 else f"gr_rot2<{R}, {rb}, {a}, {b}>")
                        out.append(f"    const double {var} = {gr}(v, w, {pauli});")
                        grads.append((var, k))
                    if kind == P.OP_ROT1:
                        out.append(f"    spd_rot1<{R}, {tb}, {a}, {pauli}, false, {rb}>(v, w, ops[{k}], trig, gacc, {ng});")
                    else:
                        out.append(f"    spd_rot2<{R}, {tb}, {a}, {b}, {pauli}, false, {rb}>(v, w, ops[{k}], trig, "
                                   f"gacc, {ng});")
                    if len(grads) == 8:
                        flush()
                else:
                    raise ValueError(kind)
            else:
                if kind == P.OP_MAT1:
                    out.append(f"    sp_mat1<{R}, {a}, {rb}>(v, ops[{k}]);")
                elif kind == P.OP_MAT2:
                    out.append(f"    sp_mat2<{R}, {a}, {b}, {rb}>(v, ops[{k}]);")
                elif kind == P.OP_CX:
                    th = "true" if a > b else "false"
                    out.append(f"    ap_cx_t<{R}, {rb}, {hi}, {lo}, {th}>(v);")
                elif kind == P.OP_CZ:
                    out.append(f"    ap_cz<{R}, {rb}, {hi}, {lo}>(v);")
                elif kind == P.OP_ROT1:
                    out.append(f"    sp_rot1<{R}, {a}, {pauli}, {rb}>(v, ops[{k}], trig);")
                elif kind == P.OP_ROT2:
                    out.append(f"    sp_rot2<{R}, {a}, {b}, {pauli}, {rb}>(v, ops[{k}], trig);")
                else:
                    raise ValueError(kind)
        flush()
    if dual:
        remap = int(p['flags']) & P.PASSF_REMAP_STORE
        if remap and ns == 1:
            out.append("    __syncthreads();")
        out.append(f"    sp_store2<{R}, {tb}, {rb}>(psi, lam, base, st[{ns if remap else ns - 1}], v, w);")
        if ng:
            out.append(f"    sp_grad_flush_wg<{tb}>(gacc, {ng}, gpart, wg);")
    elif int(p['flags']) & P.PASSF_REMAP_STORE:
        # the store permutes the tile's bits: every wave must have loaded
        # before any wave stores (with one stage no exchange barrier did it)
        if ns == 1 and store_to == "psi":
            out.append("    __syncthreads();")
        out.append(f"    sp_store<{R}, {tb}, {rb}>({store_to}, base, st[{ns}], v);")
    else:
        out.append(f"    sp_store<{R}, {tb}, {rb}>({store_to}, base, st[{ns - 1}], v);")
    return out


def _pass_kernel(R, p, stages, ops, dual, name, out_of_place=False):
    """A kernel for one pass of one program: tile index = blockIdx.x.
    out_of_place: the pass reads psi and stores into ``out`` (the last pass
    of a program that ends with a bit permutation: the permutation then reads
    ``out`` and writes psi, one sweep instead of a permutation and a copy)."""
    tb = int(p['tb'])
    lam = f"cpx<{R}> *__restrict__ lam, " if dual else ""
    dst = f"cpx<{R}> *__restrict__ out, " if out_of_place else ""
    gp = ", double *__restrict__ gpart" if dual else ""
    head = [f'extern "C" __global__ void __launch_bounds__({1 << tb}) '
            f"{name}(cpx<{R}> *__restrict__ psi, {dst}{lam}int n, const svh_pass_t *__restrict__ pass, "
            f"const svh_stage_t *__restrict__ st, const svh_op_t *__restrict__ ops, "
            f"const cpx<{R}> *__restrict__ trig{gp}) {{",
            "    const uint64_t wg = blockIdx.x;"]
    return head + _body(R, p, stages, ops, dual, "out" if out_of_place else "psi") + ["}", ""]


def stores_out_of_place(h, pi):
    """The compiled kernel of pass ``pi`` stores out of place: the last pass
    of a program that ends with a bit permutation (header off_perm set)."""
    return int(h['off_perm']) != 0 and pi == int(h['n_passes']) - 1


def batch_body(blob, pi, dtype_code, dual):
    """The pass-independent text of pass ``pi``'s batch kernel (equal texts run
    in one launch for every circuit whose pass has that structure)."""
    h, passes, stages, ops = _decode(blob)
    R = "float" if dtype_code == 0 else "double"
    return "\n".join(_body(R, passes[pi], stages, ops, dual))


# Item tables hold addresses as integers.  Converting them through the global
# address space (not straight to generic pointers) lets the compiler load the
# uniform, read-only records (pass, stages, ops, angles) with scalar loads, as
# it does for kernel-argument pointers, instead of per-lane flat loads.
_GPTR = """template <typename T>
__device__ __forceinline__ T *gptr(uint64_t a) {
    return (T *)(__attribute__((address_space(1))) T *)a;
}"""


def batch_kernel(body, dtype_code, dual, tb, name):
    """Kernel ``name`` over a batch item table (svh_batch_item_t): workgroup b
    finds its item (last wg_begin <= b) and runs ``body`` on that circuit."""
    R = "float" if dtype_code == 0 else "double"
    lam = (f"    cpx<{R}> *lam = gptr<cpx<{R}>>(it.lam);\n"
           f"    double *gpart = gptr<double>(it.gpart);\n") if dual else ""
    return (f'extern "C" __global__ void __launch_bounds__({1 << tb}) '
            f"{name}(const svh_batch_item_t *__restrict__ items, int nitems) {{\n"
            f"    const svh_batch_item_t &it = items[sp_find_item(items, nitems)];\n"
            f"    cpx<{R}> *psi = gptr<cpx<{R}>>(it.psi);\n" + lam +
            f"    const int n = it.n;\n"
            f"    const svh_pass_t *pass = gptr<const svh_pass_t>(it.pass);\n"
            f"    const svh_stage_t *st = gptr<const svh_stage_t>(it.stages) + pass->stage_begin;\n"
            f"    const svh_op_t *ops = gptr<const svh_op_t>(it.ops) + st[0].op_begin;\n"
            f"    const cpx<{R}> *trig = gptr<const cpx<{R}>>(it.trig);\n"
            f"    const uint64_t wg = (uint64_t)((int64_t)blockIdx.x - it.wg_begin);\n"
            + body + "\n}\n")


def adjoint_source(blob, dtype_code, only=None):
    """Adjoint (ψ, λ) kernels ``svh_spd_<p>`` for the tile passes of an adjoint
    program: gradient ops accumulate 2·Im⟨λ|P|ψ⟩ before applying U† to both."""
    h, passes, stages, ops = _decode(blob)
    R = "float" if dtype_code == 0 else "double"
    out = ['#include "svh_tile_ops.h"', "using namespace svh;", ""]
    for pi, p in enumerate(passes):
        if not eligible(p, dtype_code, True) or (only is not None and pi not in only):
            continue
        out += _pass_kernel(R, p, stages, ops, True, f"svh_spd_{pi}")
    return "\n".join(out)


def program_source(blob, dtype_code, only=None):
    """HIP source with one kernel per eligible tile pass of ``blob`` (restricted
    to the pass indices in ``only`` when given)."""
    h, passes, stages, ops = _decode(blob)
    R = "float" if dtype_code == 0 else "double"
    out = ['#include "svh_tile_ops.h"', "using namespace svh;", ""]
    for pi, p in enumerate(passes):
        if not eligible(p, dtype_code) or (only is not None and pi not in only):
            continue
        out += _pass_kernel(R, p, stages, ops, False, f"svh_sp_{pi}", stores_out_of_place(h, pi))
    return "\n".join(out)


def _headers_digest():
    m = hashlib.sha1()
    for f in ("svh_common.h", "svh_tile_ops.h"):
        with open(os.path.join(CSRC, f), "rb") as fh:
            m.update(fh.read())
    # the kernels use the ABI's types; the entry-point declarations do not
    # enter generated code
    with open(os.path.join(HERE, "..", "include", "qaravan_hip_types.h"), "rb") as fh:
        m.update(fh.read())
    return m.hexdigest()


def _compile_one(src, tag):
    flags = " ".join(f for f in FLAGS if not f.startswith("-I"))
    key = hashlib.sha1((src + _headers_digest() + flags).encode()).hexdigest()[:20]
    co = os.path.join(CACHE_DIR, f"sp_{key}.hsaco")
    for d in (CACHE_DIR, PACKAGE_CACHE):
        hit = os.path.join(d, f"sp_{key}.hsaco")
        if os.path.exists(hit):
            return hit
    srcf = os.path.join(CACHE_DIR, f"sp_{key}.hip")
    with open(srcf, "w") as f:
        f.write(src)
    tmp = co + f".tmp{os.getpid()}_{tag}"
    t0 = time.time()
    r = subprocess.run([HIPCC, *FLAGS, srcf, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError("specialised program compile failed:\n" + err[-4000:])
    os.replace(tmp, co)
    os.remove(srcf)
    STATS["objects_compiled"] += 1
    # a cache miss costs seconds to tens of seconds: say so, one line each
    print(f"[specialize] compiled {os.path.basename(co)} in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    return co


def compile_program(blob, dtype_code, chunks=None, dual=False):
    """Code-object paths specialising ``blob`` (compiled on first use, in
    ``chunks`` parallel hipcc runs, cached by content hash)."""
    _, passes, _, _ = _decode(blob)
    todo = [pi for pi, p in enumerate(passes) if eligible(p, dtype_code, dual)]
    if not todo:
        return []
    os.makedirs(CACHE_DIR, exist_ok=True)
    k = chunks or min(8, os.cpu_count() or 4, len(todo))
    groups = [set(todo[i::k]) for i in range(k)]
    gen = adjoint_source if dual else program_source
    srcs = [gen(blob, dtype_code, only=g) for g in groups if g]
    return _compile_all(srcs, len(srcs))


def _compile_all(srcs, workers, tag=""):
    from concurrent.futures import ThreadPoolExecutor
    t0, before = time.time(), STATS["objects_compiled"]
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        out = list(ex.map(lambda a: _compile_one(a[1], tag + str(a[0])), enumerate(srcs)))
    if STATS["objects_compiled"] > before:
        STATS["compile_wall_s"] += time.time() - t0
    return out


_LOADED = {}            # structure key -> Specialized (most recent last)
_LOADED_MAX = 64


def specialized_for(prog, dtype_code, dual=False):
    """Loaded kernels for ``prog``, shared by every program of the same pass
    structure: the kernel text holds no matrix, angle or address, so programs
    that differ only in those values (a circuit re-run with new gate
    matrices, e.g. each step of an optimisation) reuse one loaded module."""
    gen = adjoint_source if dual else program_source
    key = (dtype_code, dual, hashlib.sha1(gen(prog.blob, dtype_code).encode()).hexdigest())
    sp = _LOADED.pop(key, None)
    if sp is None:
        sp = Specialized(prog, dtype_code, dual)
        while len(_LOADED) >= _LOADED_MAX:
            _LOADED.pop(next(iter(_LOADED)))
    _LOADED[key] = sp
    return sp


class Specialized:
    """Loaded kernels of one specialised program (frees the modules on close)."""

    def __init__(self, prog, dtype_code, dual=False):
        import ctypes

        from . import _lib
        self.paths = compile_program(prog.blob, dtype_code, dual=dual)
        lib = _lib.load()
        arr = (ctypes.c_char_p * max(1, len(self.paths)))(*[p.encode() for p in self.paths])
        h = ctypes.c_void_p()
        if self.paths:
            _lib.check(lib.svh_sp_load(arr, len(self.paths), prog.n_passes, ctypes.byref(h)), "svh_sp_load")
        self.handle = h if self.paths else None
        self.n_kernels = sum(1 for p in _decode(prog.blob)[1] if eligible(p, dtype_code, dual))

    def close(self):
        if self.handle:
            import torch

            from . import _lib
            # kernels of this module may still be queued: let them finish
            # before the code object is unloaded
            torch.cuda.synchronize()
            _lib.load().svh_sp_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compile_batch(kernels, dtype_code, per_file=8, workers=8):
    """Code objects for batch kernels: ``kernels`` maps kernel name ->
    (body, dual, tb).  ``per_file`` kernels per code object (so the cache key
    of a file does not depend on the machine), ``workers`` hipcc runs at a
    time; cached by content hash like the per-pass kernels."""
    names = sorted(kernels)
    if not names:
        return []
    os.makedirs(CACHE_DIR, exist_ok=True)
    groups = [names[i:i + per_file] for i in range(0, len(names), per_file)]
    srcs = []
    for g in groups:
        if not g:
            continue
        text = ['#include "svh_tile_ops.h"', "using namespace svh;", _GPTR, ""]
        for nm in g:
            body, dual, tb = kernels[nm]
            text.append(batch_kernel(body, dtype_code, dual, tb, nm))
        srcs.append("\n".join(text))
    return _compile_all(srcs, min(workers, len(srcs)), "b")


class BatchKernels:
    """Loaded batch kernels (svh_spb_load); ``index[name]`` is the launch index."""

    def __init__(self, kernels, dtype_code):
        import ctypes

        from . import _lib
        lib = _lib.load()
        self.names = sorted(kernels)
        self.index = {nm: i for i, nm in enumerate(self.names)}
        self.paths = compile_batch(kernels, dtype_code)
        parr = (ctypes.c_char_p * len(self.paths))(*[x.encode() for x in self.paths])
        narr = (ctypes.c_char_p * len(self.names))(*[x.encode() for x in self.names])
        h = ctypes.c_void_p()
        _lib.check(lib.svh_spb_load(parr, len(self.paths), narr, len(self.names), ctypes.byref(h)), "svh_spb_load")
        self.handle = h

    def close(self):
        import ctypes

        from . import _lib
        if getattr(self, "handle", None) and self.handle.value:
            import torch
            torch.cuda.synchronize()
            _lib.load().svh_sp_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass
