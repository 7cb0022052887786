// Generic tile-pass kernels (svh_gtile.h): one compiled kernel per (dtype,
// register bits, direction) runs every pass structure, so no pass needs a
// compiler at run time.  k_gtile runs one pass of one program (a tile per
// workgroup); run_impl (svh_tile.hip) launches it for passes flagged
// SVH_PASSF_GENERIC.
#include "svh_gtile.h"

#include <mutex>
#include <unordered_map>
#include <vector>

namespace svh {

constexpr int kGtileMaxThreads = 512;

template <typename R, int RB, bool DUAL>
__global__ void __launch_bounds__(kGtileMaxThreads)
k_gtile(const cpx<R> *psi, cpx<R> *out, cpx<R> *lam, int n, const svh_pass_t *__restrict__ pass,
        const svh_stage_t *__restrict__ st, const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig,
        double *__restrict__ gpart) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *lds = reinterpret_cast<cpx<R> *>(smem_raw);
    cpx<R> *lds2 = lds + (1 << pass->L);
    gtile<R, RB, DUAL>(psi, out, lam, n, pass, st, ops, trig, gpart, blockIdx.x, lds, lds2);
}

template <typename R, int RB, bool DUAL>
static int launch_gtile(const void *psi, void *out, void *lam, int n, const svh_pass_t &hp, const svh_pass_t *dp,
                        const svh_stage_t *dst, const svh_op_t *dop, const void *trig, double *gpart,
                        hipStream_t stream) {
    auto kern = k_gtile<R, RB, DUAL>;
    if (int e = ensure_lds_limit((const void *)kern, 160 * 1024, "hipFuncSetAttribute(k_gtile)")) return e;
    const size_t lds = (DUAL ? 2 : 1) * sizeof(cpx<R>) * ((size_t)1 << hp.L);
    if (lds > 160 * 1024) return fail(-7, "generic tile pass needs more than 160 KiB of LDS");
    const uint64_t grid = 1ull << (n - hp.L);
    if (grid > 0x7fffffffull) return fail(-7, "grid too large");
    kern<<<(unsigned)grid, 1u << hp.tb, lds, stream>>>((const cpx<R> *)psi, (cpx<R> *)out, (cpx<R> *)lam, n, dp, dst,
                                                        dop, (const cpx<R> *)trig, gpart);
    return checked_launch("k_gtile");
}

// One generic pass: tile loads from psi, stores to out (psi, or the scratch
// tail for the last pass of a program that ends with a permutation).
int gtile_launch(const void *psi, void *out, void *lam, int n, int dtype, const svh_pass_t &hp,
                 const svh_pass_t *dp, const svh_stage_t *dst, const svh_op_t *dop, const void *trig,
                 double *gpart, hipStream_t stream) {
    if (hp.tb < 6 || (1 << hp.tb) > kGtileMaxThreads)
        return fail(-7, "generic tile pass needs 6 <= tb <= 9 (whole waves, at most 512 threads)");
    if (hp.flags & (SVH_PASSF_HALF_LDS | SVH_PASSF_QUARTER_LDS | SVH_PASSF_PREFETCH))
        return fail(-8, "generic tile passes take no LDS split and no prefetch");
    const bool dual = lam != nullptr;
#define SVH_GT(RT, RBV)                                                                                         \
    return dual ? launch_gtile<RT, RBV, true>(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream)       \
                : launch_gtile<RT, RBV, false>(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream)
    if (dtype == SVH_C64) {
        if (hp.rb == 3) SVH_GT(float, 3);
        if (hp.rb == 4) SVH_GT(float, 4);
    } else {
        if (hp.rb == 3) SVH_GT(double, 3);
        if (hp.rb == 4) SVH_GT(double, 4);
    }
#undef SVH_GT
    return fail(-8, "generic tile passes take 3 or 4 register bits");
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

// ------------------------------------------------ persistent batch
// k_batch_persist: a grid of resident workgroups (512 threads) takes tasks
// (svh_ptask_t) from per-queue lists; a workgroup's home queue is the XCD it
// runs on (mod nq), so a queue's circuits stay in one XCD's L2 while it has
// work, and a workgroup whose queue is drained helps the next queue.  A task first waits until its
// descriptor's dependency counter reaches dep_need (every dependency sits
// earlier in the same queue and has been taken by a running workgroup, so the
// wait ends), runs its units, makes its writes visible at device scope and
// adds its units to the `done` counter.  Waits are bounded: past the bound the
// error flag is set and the task is skipped, so the grid always drains.
constexpr int kPersistThreads = 512;
constexpr int kSpinLimit = 1 << 24;

// A descriptor field holding a device address, as a pointer into the global
// address space.  Converting through address space 1 (not straight to a
// generic pointer) lets the compiler load the uniform records behind it (pass,
// stages, ops, angles) with scalar loads, as it does for kernel arguments; from
// a generic pointer it would issue per-lane flat loads.
template <typename T>
__device__ __forceinline__ T *gptr(uint64_t a) {
    return (T *)(__attribute__((address_space(1))) T *)a;
}
// The same for records nothing writes while the kernel runs (pass, stage and
// op records, angle and term tables, slot maps): through the constant address
// space, whose loads the compiler may serve from the scalar cache even though
// the kernel stores to other global memory.
template <typename T>
__device__ __forceinline__ const T *cptr(uint64_t a) {
    return (const T *)(__attribute__((address_space(4))) const T *)a;
}

// Spin with relaxed device-scope loads (no cache maintenance per poll), then
// one device-scope acquire once the count is reached.
// Returns whether the count was not yet reached at the first poll.
__device__ __forceinline__ bool wait_counter(const int32_t *c, int need, int32_t *err) {
    int spins = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        __builtin_amdgcn_s_sleep(4);
        if (++spins > kSpinLimit) {
            atomicExch(err, 1);
            break;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return spins > 0;
}

// Fixed-order sum of v over the 512 threads (tree in LDS); valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int k = kPersistThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    return red[0];
}

struct PTerm {
    uint64_t x, z;
    double cre, cim;
};

__device__ __forceinline__ double pt_sign(uint64_t v) { return (__popcll(v) & 1) ? -1.0 : 1.0; }

// λ_i = Σ_t c_t i^{ny} (-1)^{pc((i^x)&z)} ψ[i^x] for the 1024 amplitudes of
// block blk (two per thread), and the block's Σ Re(conj(ψ_i) λ_i) into part[blk].
template <typename R>
__device__ __forceinline__ void ham_block(const svh_pdesc_t &D, int64_t blk, double *red) {
    const cpx<R> *psi = gptr<const cpx<R>>(D.psi);
    cpx<R> *lam = gptr<cpx<R>>(D.lam);
    const PTerm *terms = cptr<PTerm>(D.out);
    const uint64_t i0 = (uint64_t)blk * 1024 + threadIdx.x, i1 = i0 + 512;
    double ar0 = 0.0, ai0 = 0.0, ar1 = 0.0, ai1 = 0.0;
    for (int t = 0; t < D.ng; ++t) {
        const PTerm T = terms[t];
        const uint64_t j0 = i0 ^ T.x, j1 = i1 ^ T.x;
        const cpx<R> v0 = psi[j0], v1 = psi[j1];
        const double s0 = pt_sign(j0 & T.z), s1 = pt_sign(j1 & T.z);
        ar0 += s0 * (T.cre * v0.x - T.cim * v0.y);
        ai0 += s0 * (T.cre * v0.y + T.cim * v0.x);
        ar1 += s1 * (T.cre * v1.x - T.cim * v1.y);
        ai1 += s1 * (T.cre * v1.y + T.cim * v1.x);
    }
    lam[i0] = cpx<R>{(R)ar0, (R)ai0};
    lam[i1] = cpx<R>{(R)ar1, (R)ai1};
    const cpx<R> p0 = psi[i0], p1 = psi[i1];
    const double e = ((double)p0.x * ar0 + (double)p0.y * ai0) + ((double)p1.x * ar1 + (double)p1.y * ai1);
    const double tot = block_sum(e, red);
    if (threadIdx.x == 0) gptr<double>(D.part)[blk] = tot;
}

// Slot s of a reduction: Σ_r part[r·ng + s] in a fixed order.
__device__ __forceinline__ void red_slot(const svh_pdesc_t &D, int s, double *red) {
    const double *part = gptr<const double>(D.part);
    double acc = 0.0;
    for (int r = threadIdx.x; r < D.rows; r += kPersistThreads) acc += part[(int64_t)r * D.ng + s];
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) {
        if (D.slotsum) gptr<double>(D.slotsum)[s] = tot;
        else gptr<double>(D.out)[cptr<int32_t>(D.slot_param)[s]] = tot;
    }
}

// complex64: two resident workgroups per CU = 4 waves per SIMD (HIP's second
// bound is waves per execution unit), at most 128 VGPRs per lane; complex128
// needs more registers and runs one workgroup per CU
template <typename R>
__global__ void __launch_bounds__(kPersistThreads, sizeof(R) == 4 ? 4 : 2)
k_batch_persist(const svh_pdesc_t *__restrict__ descs, const svh_ptask_t *__restrict__ tasks,
                const int32_t *__restrict__ qbeg, int nq, int32_t *heads, int32_t *ctr, int32_t *err,
                int xcd_local, unsigned long long *stats) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ int s_task;
    cpx<R> *lds = reinterpret_cast<cpx<R> *>(smem_raw);
    double *red = reinterpret_cast<double *>(smem_raw);
    // home queue: the XCD this workgroup runs on (hardware register XCC_ID,
    // bits 3:0), so a queue's circuits stay in one L2
    const int xcc = (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15);
    // XCD-local mode: a workgroup serves only its own XCD's queue (no help
    // across XCDs), so every dependency stays inside one L2
    int q = xcc % nq, drained = (xcd_local && xcc >= nq) ? nq : 0;  // used by thread 0
    const int reach = xcd_local ? 1 : nq;
    // thread 0's clock accounting (100 MHz ticks): dependency waits, everything
    // from a task's claim to its counter update, tasks, waits that had to spin
    unsigned long long t_wait = 0, t_task = 0, n_task = 0, n_spun = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int t = -1;
            while (drained < reach) {
                const int k = atomicAdd(&heads[q], 1);
                if (qbeg[q] + k < qbeg[q + 1]) {
                    t = qbeg[q] + k;
                    break;
                }
                q = (q + 1) % nq;  // this queue is drained: help the next one
                ++drained;
            }
            s_task = t;
        }
        __syncthreads();
        // the same value in every lane: say so, so that the task record, its
        // descriptor and every record they point to load through the scalar unit
        const int t = __builtin_amdgcn_readfirstlane(s_task);
        __syncthreads();  // every thread has read s_task before thread 0 writes it again
        if (t < 0) break;
        const svh_ptask_t T = tasks[t];
        const svh_pdesc_t &D = descs[T.desc];
        const unsigned long long t0 = stats && threadIdx.x == 0 ? wall_clock64() : 0;
        if (D.dep >= 0) {
            if (threadIdx.x == 0) {
                const bool spun = wait_counter(ctr + D.dep, D.dep_need, err);
                if (stats) {
                    const unsigned long long t1 = wall_clock64();
                    t_wait += t1 - t0;
                    n_spun += spun ? 1 : 0;
                }
            }
            __syncthreads();
        }
        for (int u = T.first; u < T.first + T.count; ++u) {
            if (u > T.first) __syncthreads();  // the previous unit is done with LDS
            if (D.kind == SVH_PT_FWD || D.kind == SVH_PT_ADJ) {
                const auto *pass = cptr<svh_pass_t>(D.pass);
                cpx<R> *psi = gptr<cpx<R>>(D.psi);
                cpx<R> *lds2 = lds + (1 << pass->L);
                if (D.kind == SVH_PT_ADJ)
                    gtile<R, 3, true>(psi, psi, gptr<cpx<R>>(D.lam), D.n, pass,
                                      cptr<svh_stage_t>(D.stages),
                                      cptr<svh_op_t>(D.ops),
                                      cptr<cpx<R>>(D.trig), gptr<double>(D.part),
                                      (uint64_t)u, lds, lds2);
                else
                    gtile<R, 3, false>(psi, psi, nullptr, D.n, pass, cptr<svh_stage_t>(D.stages),
                                       cptr<svh_op_t>(D.ops),
                                       cptr<cpx<R>>(D.trig), nullptr, (uint64_t)u, lds, lds2);
            } else if (D.kind == SVH_PT_HAM) {
                ham_block<R>(D, u, red);
            } else if (D.kind == SVH_PT_RED) {
                red_slot(D, u, red);
            } else if (D.kind == SVH_PT_SCT) {
                if (threadIdx.x == 0) {
                    const double *ss = gptr<const double>(D.slotsum);
                    const int32_t *sp = cptr<int32_t>(D.slot_param);
                    double *out = gptr<double>(D.out);
                    for (int s = 0; s < D.ng; ++s) out[sp[s]] += ss[s];
                }
            }
        }
        // every wave makes its writes visible before the counters say the
        // units are done: device-wide (write the XCD's L2 back) or, in
        // XCD-local mode, where every reader shares this XCD's L2, by waiting
        // until the stores have reached it (L1 writes through)
        if (xcd_local) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            if (D.done >= 0) __hip_atomic_fetch_add(ctr + D.done, T.count, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (D.fin >= 0) {
                const int before =
                    __hip_atomic_fetch_add(ctr + D.fin, T.count, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                if (before + T.count == D.fin_need && D.tstamp)
                    *gptr<uint64_t>(D.tstamp) = wall_clock64();
            } else if (D.tstamp) {
                *gptr<uint64_t>(D.tstamp) = wall_clock64();
            }
            if (stats) {
                t_task += wall_clock64() - t0;
                ++n_task;
            }
        }
    }
    if (stats && threadIdx.x == 0) {
        atomicAdd(stats + 0, t_wait);
        atomicAdd(stats + 1, t_task);
        atomicAdd(stats + 2, n_task);
        atomicAdd(stats + 3, n_spun);
        atomicAdd(stats + 4, 1ull);
    }
}

// XCD of every workgroup of a grid (the persistent grid's shape and LDS), so
// the host knows whether each of the nq queues has workgroups on its own XCD.
__global__ void __launch_bounds__(kPersistThreads) k_xcc_census(int32_t *out) {
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

static int persist_grid(int dtype, size_t lds, int *grid, uint32_t *xcc_mask) {
    static std::mutex mu;
    static std::unordered_map<uint64_t, PersistShape> cache;
    int dev = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    const uint64_t key = ((uint64_t)dev << 40) | ((uint64_t)dtype << 32) | (uint64_t)lds;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *grid = it->second.grid;
        *xcc_mask = it->second.xcc_mask;
        return 0;
    }
    const void *kern = dtype == SVH_C64 ? (const void *)k_batch_persist<float> : (const void *)k_batch_persist<double>;
    // the dynamic LDS this launch asks for (the kernel also holds a few static bytes)
    if (int e = hip_status(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                           "hipFuncSetAttribute(k_batch_persist)"))
        return e;
    int per_cu = 0, cus = 0;
    if (int e = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kPersistThreads, lds),
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
    if (hipStreamIsCapturing(nullptr, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
        int32_t *d = nullptr;
        if (hipMalloc(&d, sizeof(int32_t) * (size_t)*grid) == hipSuccess) {
            std::vector<int32_t> h((size_t)*grid, -1);
            k_xcc_census<<<*grid, kPersistThreads, lds, 0>>>(d);
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
                                    int dtype, size_t lds_bytes, int max_workgroups, int xcd_local, int *used_mode,
                                    uint64_t *stats_dev, void *stream) {
    if (!descs_dev || !tasks_dev || !qbeg_dev || !heads_dev || !err_dev || (ncounters > 0 && !counters_dev))
        return fail(-1, "NULL pointer");
    if (nq < 1 || ncounters < 0) return fail(-4, "bad queue or counter count");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "bad dtype");
    if (lds_bytes > 159 * 1024) return fail(-7, "persistent batch needs more than 159 KiB of LDS");
    const size_t need = lds_bytes > kPersistThreads * sizeof(double) ? lds_bytes : kPersistThreads * sizeof(double);
    int grid = 0;
    uint32_t xcc_mask = 0;
    if (int e = persist_grid(dtype, need, &grid, &xcc_mask)) return e;
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
    if (dtype == SVH_C64)
        k_batch_persist<float><<<grid, kPersistThreads, need, st>>>(D, T, qbeg_dev, nq, heads_dev, counters_dev, err_dev,
                                                                    mode, (unsigned long long *)stats_dev);
    else
        k_batch_persist<double><<<grid, kPersistThreads, need, st>>>(D, T, qbeg_dev, nq, heads_dev, counters_dev,
                                                                     err_dev, mode, (unsigned long long *)stats_dev);
    return checked_launch("k_batch_persist");
}
