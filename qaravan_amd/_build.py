"""Compile the HIP sources under csrc/ into libqaravan_hip.so (gfx950).

Each .hip file is compiled to an object in parallel, then linked into one
shared library next to this file.  Only hipcc is needed; no torch extension
machinery is involved, so the library exposes a plain C ABI.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqaravan_hip.so")
ARCH = os.environ.get("QARAVAN_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # no packed FP32 (v_pk_*_f32): plain scalar FP32 throughout
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# per-file flags.  The generic tile kernels keep their register arrays in
# VGPRs only if LLVM does not merge look-alike switch arms into one block that
# addresses the array through a phi of pointers (SimplifyCFG sinking), which
# would force the array into scratch memory.
GTILE_FLAGS = ["-mllvm", "-simplifycfg-sink-common=false"]


def file_flags(name):
    """Extra flags of one source: the generic tile kernels (svh_gtile.hip and
    its per-instantiation files svh_gt_*.hip, svh_pb_*.hip)."""
    return GTILE_FLAGS if name == "svh_gtile.hip" or name.startswith(("svh_gt_", "svh_pb_")) else []


def _compile(src, objdir):
    obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *FLAGS, *file_flags(os.path.basename(src)), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj


def _up_to_date(src, objdir):
    """The object of ``src`` is newer than the source and every header it
    may include (csrc/*.h and include/*.h)."""
    obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
    if not os.path.exists(obj):
        return False
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(HERE, "..", "include")
    deps += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return os.path.getmtime(obj) > max(os.path.getmtime(d) for d in deps)


def build(verbose=True, incremental=False):
    """Compile and link the library.  ``incremental``: reuse objects that are
    newer than their sources and headers (a development shortcut; the default
    recompiles everything)."""
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    todo = [s for s in srcs if not (incremental and _up_to_date(s, objdir))]
    with ThreadPoolExecutor(max_workers=max(1, min(8, len(todo)))) as ex:
        list(ex.map(lambda s: _compile(s, objdir), todo))
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
