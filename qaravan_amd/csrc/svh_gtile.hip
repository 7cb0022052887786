// Generic tile-pass kernels (svh_gtile_kern.h): one compiled kernel per (dtype,
// register bits, direction) runs every pass structure, so no pass needs a
// compiler at run time.  k_gtile runs one pass of one program (a tile per
// workgroup); run_impl (svh_tile.hip) launches it for passes flagged
// SVH_PASSF_GENERIC.  This file holds the host side: the dispatch to the
// per-instantiation files, the pattern check and the persistent batch's
// launch (k_batch_persist, svh_pb_*.hip).
#include "svh_gtile.h"
#include "svh_gtile_entry.h"

#include <mutex>
#include <unordered_map>
#include <vector>

namespace svh {

constexpr int kGtileMaxThreads = 512;

// One generic pass: tile loads from psi, stores to out (psi, or the scratch
// tail for the last pass of a program that ends with a permutation).
int gtile_launch(const void *psi, void *out, void *lam, int n, int dtype, const svh_pass_t &hp,
                 const svh_pass_t *dp, const svh_stage_t *dst, const svh_op_t *dop, const void *trig,
                 double *gpart, hipStream_t stream) {
    if (hp.tb < 6 || (1 << hp.tb) > kGtileMaxThreads)
        return fail(-7, "generic tile pass needs 6 <= tb <= 9 (whole waves, at most 512 threads)");
    if (hp.flags & (SVH_PASSF_HALF_LDS | SVH_PASSF_QUARTER_LDS | SVH_PASSF_PREFETCH))
        return fail(-8, "generic tile passes take no LDS split and no prefetch");
    if (hp.rb != 3 && hp.rb != 4) return fail(-8, "generic tile passes take 3 or 4 register bits");
    if (dtype == SVH_C64)
        return hp.rb == 3 ? gtile_launch_f3(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream)
                          : gtile_launch_f4(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream);
    return hp.rb == 3 ? gtile_launch_d3(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream)
                      : gtile_launch_d4(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream);
}

// Host check: stage pattern pid of the (rb, dual) library is exactly the op
// sequence ops[0..nops) (kind, register bits, Pauli code, gradient flag).
bool gtile_pattern_matches(int rb, bool dual, int pid, const svh_op_t *ops, int nops) {
    int count = 0;
    const int *off = nullptr;
    const signed char(*tab)[5] = nullptr;
    if (rb == 3 && !dual) count = kPatCount3f, off = kPatOff3f, tab = kPatOps3f;
    else if (rb == 3 && dual) count = kPatCount3d, off = kPatOff3d, tab = kPatOps3d;
    else if (rb == 4 && !dual) count = kPatCount4f, off = kPatOff4f, tab = kPatOps4f;
    else if (rb == 4 && dual) count = kPatCount4d, off = kPatOff4d, tab = kPatOps4d;
    if (pid < 0 || pid >= count || off[pid + 1] - off[pid] != nops) return false;
    for (int k = 0; k < nops; ++k) {
        const signed char *t = tab[off[pid] + k];
        const svh_op_t &o = ops[k];
        const int grad = (dual && (o.flags & SVH_OPF_GRAD)) ? 1 : 0;
        if (t[0] != o.kind || t[1] != o.a || t[2] != o.b || t[3] != o.pauli || t[4] != grad) return false;
    }
    return true;
}

// XCD of every workgroup of a grid (the persistent grid's shape and LDS), so
// the host knows whether each of the nq queues has workgroups on its own XCD.
__global__ void __launch_bounds__(512) k_xcc_census(int32_t *out) {
    extern __shared__ unsigned char smem_raw[];
    if (threadIdx.x == 0) {
        smem_raw[0] = 0;
        out[blockIdx.x] = (int32_t)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15);
    }
}

__global__ void k_fill_i32(int32_t *p, int64_t count, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Resident workgroups of the persistent kernel on the current device (the
// occupancy of one dtype and LDS size), computed once per (device, dtype, lds).
struct PersistShape {
    int grid;
    uint32_t xcc_mask;  // XCDs the grid's workgroups ran on (census), 0 = unknown
};

// lds: the kernel's dynamic LDS; census_lds: the LDS the census kernel takes
// so that its workgroups are resident in the same numbers per CU.
static int persist_grid(const void *kern, int threads, size_t lds, size_t census_lds, int *grid,
                        uint32_t *xcc_mask) {
    static std::mutex mu;
    static std::unordered_map<uint64_t, PersistShape> cache;
    int dev = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    const uint64_t key = ((uint64_t)dev << 48) ^ ((uint64_t)(uintptr_t)kern << 20) ^ (uint64_t)lds;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *grid = it->second.grid;
        *xcc_mask = it->second.xcc_mask;
        return 0;
    }
    // the dynamic LDS this launch asks for (the kernel also holds a few static bytes)
    if (int e = hip_status(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                           "hipFuncSetAttribute(k_batch_persist)"))
        return e;
    int per_cu = 0, cus = 0;
    if (int e = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds),
                           "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return e;
    if (int e = hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                           "hipDeviceGetAttribute"))
        return e;
    if (per_cu < 1) return fail(-7, "persistent batch kernel does not fit on a CU");
    *grid = per_cu * cus;
    // census: which XCDs a grid of this shape occupies (synchronous, once per
    // device and shape; skipped while a graph is being captured)
    uint32_t mask = 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(nullptr, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone &&
        ensure_lds_limit((const void *)k_xcc_census, 160 * 1024, "hipFuncSetAttribute(k_xcc_census)") == 0) {
        int32_t *d = nullptr;
        if (hipMalloc(&d, sizeof(int32_t) * (size_t)*grid) == hipSuccess) {
            std::vector<int32_t> h((size_t)*grid, -1);
            k_xcc_census<<<*grid, threads, census_lds, 0>>>(d);
            if (hipMemcpy(h.data(), d, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost) == hipSuccess)
                for (int32_t x : h)
                    if (x >= 0 && x < 32) mask |= 1u << x;
            (void)hipFree(d);
        }
        (void)hipGetLastError();
    }
    *xcc_mask = mask;
    cache[key] = PersistShape{*grid, mask};
    return 0;
}

}  // namespace svh

using namespace svh;

extern "C" int svh_batch_persistent(const void *descs_dev, const void *tasks_dev, const int32_t *qbeg_dev, int nq,
                                    int32_t *heads_dev, int32_t *counters_dev, int ncounters, int32_t *err_dev,
                                    int dtype, size_t lds_bytes, int threads, int max_workgroups, int xcd_local,
                                    int *used_mode, uint64_t *stats_dev, void *stream) {
    if (!descs_dev || !tasks_dev || !qbeg_dev || !heads_dev || !err_dev || (ncounters > 0 && !counters_dev))
        return fail(-1, "NULL pointer");
    if (nq < 1 || ncounters < 0) return fail(-4, "bad queue or counter count");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "bad dtype");
    // workgroup size: 512 threads (two resident per CU), or for complex64 256
    // (four per CU, tiles of at most 2^11 amplitudes)
    if (threads == 0) threads = 512;
    if (threads != 512 && !(threads == 256 && dtype == SVH_C64))
        return fail(-4, "persistent batch: workgroups of 512 threads, or 256 for complex64");
    const bool small = threads == 256;
    // the kernel's static LDS holds the ψ and λ tiles of an adjoint task (λ
    // at its fixed offset of threads·8 complex64 or threads·4 complex128
    // amplitudes), the reduction tree and the task index; no dynamic LDS
    const size_t lam_bytes = dtype == SVH_C64 ? (size_t)threads * 8 * 8 : (size_t)threads * 4 * 16;
    if (lds_bytes > 2 * lam_bytes)
        return fail(-7, "persistent batch: tiles larger than the workgroup's LDS (λ offset)");
    const size_t dyn_lds = 0;
    const void *kern = dtype == SVH_C128 ? persist_kernel_d3() : small ? persist_kernel_f3_256() : persist_kernel_f3();
    int grid = 0;
    uint32_t xcc_mask = 0;
    if (int e = persist_grid(kern, threads, dyn_lds, 2 * lam_bytes, &grid, &xcc_mask)) return e;
    if (max_workgroups > 0 && grid > max_workgroups) grid = max_workgroups;
    // XCD-local synchronisation only when the full grid covers exactly the
    // XCDs 0..nq-1, so every queue is served from inside its own XCD
    const int mode = (xcd_local && max_workgroups <= 0 && nq <= 32 && xcc_mask == ((1ull << nq) - 1)) ? 1 : 0;
    if (used_mode) *used_mode = mode;
    hipStream_t st = (hipStream_t)stream;
    // counters, queue heads and the error flag start at zero (fill kernels,
    // which a captured graph replays correctly)
    if (ncounters > 0) k_fill_i32<<<64, 256, 0, st>>>(counters_dev, ncounters, 0);
    k_fill_i32<<<1, 64, 0, st>>>(heads_dev, nq, 0);
    k_fill_i32<<<1, 64, 0, st>>>(err_dev, 1, 0);
    if (int e = checked_launch("k_fill_i32")) return e;
    const auto *D = static_cast<const svh_pdesc_t *>(descs_dev);
    const auto *T = static_cast<const svh_ptask_t *>(tasks_dev);
    auto *S = (unsigned long long *)stats_dev;
    auto launch = dtype == SVH_C128 ? persist_launch_d3 : small ? persist_launch_f3_256 : persist_launch_f3;
    return launch((unsigned)grid, (unsigned)threads, dyn_lds, st, D, T, qbeg_dev, nq, heads_dev, counters_dev,
                  err_dev, mode, S);
}
