// Kernel templates of the generic tile passes (k_gtile) and of the persistent
// batch (k_batch_persist).  Each (dtype, register bits) instantiation is
// compiled in a file of its own (svh_gt_*.hip, svh_pb_*.hip) so the large
// pattern library compiles in parallel; svh_gtile.hip dispatches to them.
#pragma once
#include "svh_gtile.h"
#include "svh_gtile_entry.h"

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

// ------------------------------------------------ persistent batch
// k_batch_persist: a grid of resident workgroups (512 threads, or 256) takes tasks
// (svh_ptask_t) from per-queue lists; a workgroup's home queue is the XCD it
// runs on (mod nq), so a queue's circuits stay in one XCD's L2 while it has
// work, and a workgroup whose queue is drained helps the next queue.  A task first waits until its
// descriptor's dependency counter reaches dep_need (every dependency sits
// earlier in the same queue and has been taken by a running workgroup, so the
// wait ends), runs its units, makes its writes visible at device scope and
// adds its units to the `done` counter.  Waits are bounded: past the bound the
// error flag is set and the task is skipped, so the grid always drains.
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

// Fixed-order sum of v over the TH threads (tree in LDS); valid in thread 0.
template <int TH>
__device__ __forceinline__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int k = TH / 2; k > 0; k >>= 1) {
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
// block blk (1024 / TH per thread, TH apart), and the block's Σ Re(conj(ψ_i) λ_i)
// into part[blk].
template <typename R, int TH>
__device__ __forceinline__ void ham_block(const svh_pdesc_t &D, int64_t blk, double *red) {
    constexpr int A = 1024 / TH;
    const cpx<R> *psi = gptr<const cpx<R>>(D.psi);
    cpx<R> *lam = gptr<cpx<R>>(D.lam);
    const PTerm *terms = cptr<PTerm>(D.out);
    const uint64_t i0 = (uint64_t)blk * 1024 + threadIdx.x;
    double ar[A], ai[A];
#pragma unroll
    for (int a = 0; a < A; ++a) ar[a] = ai[a] = 0.0;
    for (int t = 0; t < D.ng; ++t) {
        const PTerm T = terms[t];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const uint64_t j = (i0 + (uint64_t)a * TH) ^ T.x;
            const cpx<R> v = psi[j];
            const double sg = pt_sign(j & T.z);
            ar[a] += sg * (T.cre * v.x - T.cim * v.y);
            ai[a] += sg * (T.cre * v.y + T.cim * v.x);
        }
    }
    double e = 0.0;
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const uint64_t i = i0 + (uint64_t)a * TH;
        lam[i] = cpx<R>{(R)ar[a], (R)ai[a]};
        const cpx<R> p = psi[i];
        e += (double)p.x * ar[a] + (double)p.y * ai[a];
    }
    const double tot = block_sum<TH>(e, red);
    if (threadIdx.x == 0) gptr<double>(D.part)[blk] = tot;
}

// Slot s of a reduction: Σ_r part[r·ng + s] in a fixed order.
template <int TH>
__device__ __forceinline__ void red_slot(const svh_pdesc_t &D, int s, double *red) {
    const double *part = gptr<const double>(D.part);
    double acc = 0.0;
    for (int r = threadIdx.x; r < D.rows; r += TH) acc += part[(int64_t)r * D.ng + s];
    const double tot = block_sum<TH>(acc, red);
    if (threadIdx.x == 0) {
        if (D.slotsum) gptr<double>(D.slotsum)[s] = tot;
        else gptr<double>(D.out)[cptr<int32_t>(D.slot_param)[s]] = tot;
    }
}

// Workgroups of TH threads running tiles of RB register bits (the library
// builds TH = 512, RB = 3).  complex64: two resident workgroups per CU = 4
// waves per SIMD (HIP's second bound is waves per execution unit), at most 128
// VGPRs per lane; complex128 needs more registers and runs one workgroup per CU.
template <typename R, int TH, int RB>
constexpr int persist_waves_per_eu() {
    return sizeof(R) == 4 ? 4 : 2;
}

// The λ tile's place in LDS: right after the largest ψ tile the batch plans
// for this workgroup size (2^12 complex64 amplitudes with 512 threads, 2^11
// with 256, and 2^11 complex128), a constant the LDS instructions take as
// their offset (at most 32 KiB).
template <typename R, int TH, int RB>
constexpr int persist_lam_offset() {
    return sizeof(R) == 4 ? (TH << RB) : (TH << (RB - 1));
}

template <typename R, int TH, int RB>
__global__ void __launch_bounds__(TH, (persist_waves_per_eu<R, TH, RB>()))
k_batch_persist(const svh_pdesc_t *__restrict__ descs, const svh_ptask_t *__restrict__ tasks,
                const int32_t *__restrict__ qbeg, int nq, int32_t *heads, int32_t *ctr, int32_t *err,
                int xcd_local, unsigned long long *stats) {
    // One static LDS array, at a compile-time address, so that the tile
    // slots' addresses fold into the LDS instructions: the ψ tile, the λ tile
    // at its fixed offset, and over the same bytes the reduction tree and the
    // claimed task index (used only between barriers where no tile is live).
    constexpr int LOFF = persist_lam_offset<R, TH, RB>();
    __shared__ cpx<R> tiles[2 * LOFF];
    cpx<R> *lds = tiles;
    double *red = reinterpret_cast<double *>(tiles);
    int &s_task = *reinterpret_cast<int *>(tiles);
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
                // a tile wider than the workgroup, with other register bits, or
                // larger than the λ offset: refuse it
                if ((1 << pass->tb) > TH || pass->rb != RB || (1 << pass->L) > persist_lam_offset<R, TH, RB>()) {
                    if (threadIdx.x == 0) atomicExch(err, 2);
                    break;
                }
                cpx<R> *psi = gptr<cpx<R>>(D.psi);
                if (D.kind == SVH_PT_ADJ)
                    gtile<R, RB, true, LOFF>(psi, psi, gptr<cpx<R>>(D.lam), D.n, pass,
                                             cptr<svh_stage_t>(D.stages),
                                             cptr<svh_op_t>(D.ops),
                                             cptr<cpx<R>>(D.trig), gptr<double>(D.part),
                                             (uint64_t)u, lds, nullptr);
                else
                    gtile<R, RB, false>(psi, psi, nullptr, D.n, pass, cptr<svh_stage_t>(D.stages),
                                       cptr<svh_op_t>(D.ops),
                                       cptr<cpx<R>>(D.trig), nullptr, (uint64_t)u, lds, nullptr);
            } else if (D.kind == SVH_PT_HAM) {
                ham_block<R, TH>(D, u, red);
            } else if (D.kind == SVH_PT_RED) {
                red_slot<TH>(D, u, red);
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

}  // namespace svh
