// Fused tile passes: one launch applies a run of gates to every tile of 2^L
// amplitudes (L = TB + RB).  A workgroup of 2^TB threads owns one tile; each
// thread holds 2^RB amplitudes in registers.  A "stage" fixes which RB
// tile-local bits live in the register index; every gate of the stage acts on
// those bits only, so it is applied in registers with compile-time indexing.
// Between stages the tile is re-distributed through LDS.  The first stage's
// layout is used for the HBM load and the last one's for the store; the host
// scheduler keeps the low tile bits on lanes there so each wave moves
// contiguous 512-byte (c64) / 1 KiB (c128) runs.
//
// DUAL = true is the adjoint sweep: the same walk over ψ and λ together, with
// 2·Im⟨λ|P|ψ⟩ accumulated in FP64 for every gate flagged SVH_OPF_GRAD before
// U† is applied to both vectors.
#include "svh_tile_ops.h"

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace svh {

constexpr int kPermTileBits = 11;  // bit-permutation tiles: 2048 amplitudes
constexpr int kMaxFinalCx = 8;     // CNOTs a program's final permutation may apply

// ------------------------------------------------ the pass kernel
// One tile of one pass: workgroup-local work shared by k_tile (one circuit,
// tile = blockIdx.x) and k_tile_batch (many circuits in one launch).
template <typename R, int TB, int RB, bool DUAL>
__device__ __forceinline__ void tile_body(cpx<R> *__restrict__ psi, cpx<R> *__restrict__ lam, int n,
                                          const svh_pass_t *__restrict__ pass,
                                          const svh_stage_t *__restrict__ stages,
                                          const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig,
                                          double *__restrict__ gpart, uint64_t wg) {
    static_assert(RB == 4, "stage tables hold 16 registers");
    constexpr int L = TB + RB;
    constexpr int NR = 1 << RB;
    constexpr int NW = ((1 << TB) + 63) / 64;  // waves per workgroup
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *lds = reinterpret_cast<cpx<R> *>(smem_raw);
    cpx<R> *lds2 = lds + (1 << L);
    double *gacc = reinterpret_cast<double *>(smem_raw + (DUAL ? 2 : 1) * sizeof(cpx<R>) * (1 << L));

    const int tid = threadIdx.x;
    const int n_stages = pass->n_stages;
    const int s0 = pass->stage_begin;
    const int ng = pass->n_grad;
    // Tile base address: workgroup id deposited into the non-tile bits.
    uint64_t tmask = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) tmask |= 1ull << pass->bits[i];
    uint64_t base = 0;
    {
        uint64_t g = wg;
        for (int b = 0; b < n && g; ++b) {
            if (!((tmask >> b) & 1)) {
                base |= (g & 1ull) << b;
                g >>= 1;
            }
        }
    }

    cpx<R> v[NR];
    cpx<R> w[DUAL ? NR : 1];

    // ---- load in stage 0's layout
    {
        const svh_stage_t &S = stages[s0];
        const cpx<R> *p = psi + (base | thr_global<TB>(S, tid));
        const cpx<R> *q = DUAL ? lam + (base | thr_global<TB>(S, tid)) : nullptr;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            v[j] = p[S.goff[j]];
            if constexpr (DUAL) w[j] = q[S.goff[j]];
        }
    }

    for (int s = 0; s < n_stages; ++s) {
        const svh_stage_t &S = stages[s0 + s];
        if (s > 0) {
            // ---- re-distribute the tile from stage s-1's layout to stage s's
            const svh_stage_t &P = stages[s0 + s - 1];
            const int tp = lds_swz(thr_local(P, tid));
            if (s > 1) __syncthreads();  // previous exchange's reads are done
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                const int li = tp ^ P.lsw[j];
                lds[li] = v[j];
                if constexpr (DUAL) lds2[li] = w[j];
            }
            __syncthreads();
            const int tn = lds_swz(thr_local(S, tid));
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                const int li = tn ^ S.lsw[j];
                v[j] = lds[li];
                if constexpr (DUAL) w[j] = lds2[li];
            }
        }
        // ---- gates of this stage, in registers
        for (int o = S.op_begin; o < S.op_end; ++o) {
            const svh_op_t &op = ops[o];
            double g = 0.0;
            bool has_g = false;
            apply_op<R, RB, DUAL>(v, w, op, trig, &g, &has_g);
            if constexpr (DUAL) {
                if (has_g) {
                    g = wave_sum<((1 << TB) < 64 ? (1 << TB) : 64)>(g);
                    if ((tid & 63) == 0) gacc[(tid >> 6) * ng + op.grad_slot] = 2.0 * g;
                }
            }
        }
    }

    // ---- store in the last stage's layout
    {
        const svh_stage_t &S = stages[s0 + n_stages - 1];
        cpx<R> *p = psi + (base | thr_global<TB>(S, tid));
        cpx<R> *q = DUAL ? lam + (base | thr_global<TB>(S, tid)) : nullptr;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            p[S.goff[j]] = v[j];
            if constexpr (DUAL) q[S.goff[j]] = w[j];
        }
    }

    if constexpr (DUAL) {
        if (ng > 0) {
            __syncthreads();
            for (int sl = tid; sl < ng; sl += (1 << TB)) {
                double acc = 0.0;
                for (int wv = 0; wv < NW; ++wv) acc += gacc[wv * ng + sl];
                gpart[wg * ng + sl] = acc;
            }
        }
    }
}

template <typename R, int TB, int RB, bool DUAL>
__global__ void __launch_bounds__(1 << TB)
k_tile(cpx<R> *__restrict__ psi, cpx<R> *__restrict__ lam, int n,
       const svh_pass_t *__restrict__ pass, const svh_stage_t *__restrict__ stages,
       const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig,
       double *__restrict__ gpart) {
    tile_body<R, TB, RB, DUAL>(psi, lam, n, pass, stages, ops, trig, gpart, blockIdx.x);
}

// Many circuits' passes in one launch: workgroup b belongs to the item whose
// [wg_begin, next wg_begin) range holds b (items sorted by wg_begin).
template <typename R, int TB, int RB, bool DUAL>
__global__ void __launch_bounds__(1 << TB)
k_tile_batch(const svh_batch_item_t *__restrict__ items, int nitems) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {  // last item with wg_begin <= b
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].wg_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    const svh_batch_item_t &it = items[lo];
    tile_body<R, TB, RB, DUAL>(reinterpret_cast<cpx<R> *>(it.psi), reinterpret_cast<cpx<R> *>(it.lam), it.n,
                               reinterpret_cast<const svh_pass_t *>(it.pass),
                               reinterpret_cast<const svh_stage_t *>(it.stages),
                               reinterpret_cast<const svh_op_t *>(it.ops),
                               reinterpret_cast<const cpx<R> *>(it.trig), reinterpret_cast<double *>(it.gpart),
                               (uint64_t)(b - it.wg_begin));
}

// ------------------------------------------------ prefetching forward pass
// Persistent variant of k_tile for forward passes: each workgroup walks tiles
// t = blockIdx.x, blockIdx.x + gridDim.x, ...  While tile t is computed in
// registers (exchanges through the swizzled buffer `lds`), tile t + gridDim.x
// is already streaming HBM -> LDS into `pf` with global_load_lds_dwordx4 (16 B
// per lane, no VGPRs held), so the memory system stays busy during the ALU
// work.  Stage 0 reads its registers straight from `pf` (tile-local order).
template <typename R, int TB>
__global__ void __launch_bounds__(1 << TB)
k_tile_pf(cpx<R> *__restrict__ psi, int n, const svh_pass_t *__restrict__ pass,
          const svh_stage_t *__restrict__ stages, const svh_op_t *__restrict__ ops,
          const cpx<R> *__restrict__ trig, uint64_t ntiles) {
    constexpr int RB = 4;
    constexpr int L = TB + RB;
    constexpr int NR = 1 << RB;
    constexpr int NT = 1 << TB;
    constexpr int CH = 16 / (int)sizeof(cpx<R>);  // amplitudes per 16-byte chunk
    constexpr int NCH = (1 << L) / CH / NT;        // chunks per thread per tile
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *pf = reinterpret_cast<cpx<R> *>(smem_raw);
    cpx<R> *lds = pf + (1 << L);

    const int tid = threadIdx.x;
    const int n_stages = pass->n_stages;
    const int s0 = pass->stage_begin;
    uint64_t tmask = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) tmask |= 1ull << pass->bits[i];
    // global offset of tile element li (deposit of li's bits into the tile bits)
    auto gdep = [&](uint64_t li) {
        uint64_t a = 0;
#pragma unroll
        for (int p = 0; p < L; ++p) a |= ((li >> p) & 1ull) << pass->bits[p];
        return a;
    };
    auto tile_base = [&](uint64_t t) {
        uint64_t base = 0, g = t;
        for (int b = 0; b < n && g; ++b) {
            if (!((tmask >> b) & 1)) {
                base |= (g & 1ull) << b;
                g >>= 1;
            }
        }
        return base;
    };
    const uint64_t gthr = gdep((uint64_t)CH * tid);
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    auto prefetch = [&](uint64_t t) {
        const cpx<R> *src = psi + (tile_base(t) | gthr);
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
            const uint64_t gk = gdep((uint64_t)CH * k * NT);
            char *dst = reinterpret_cast<char *>(pf) + 16 * (k * NT + 64 * wv);
            __builtin_amdgcn_global_load_lds((const void *)(src + gk),
                                             (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
        }
    };

    cpx<R> v[NR];
    cpx<R> w[1];
    uint64_t t = blockIdx.x;
    if (t < ntiles) prefetch(t);
    for (; t < ntiles; t += gridDim.x) {
        const uint64_t base = tile_base(t);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // tile t is in pf for every wave
        {
            const svh_stage_t &S = stages[s0];
            const int tl = thr_local(S, tid);
#pragma unroll
            for (int j = 0; j < NR; ++j) v[j] = pf[tl | lds_swz(S.lsw[j])];
        }
        __syncthreads();  // everyone has read pf
        if (t + gridDim.x < ntiles) prefetch(t + gridDim.x);

        for (int s = 0; s < n_stages; ++s) {
            const svh_stage_t &S = stages[s0 + s];
            if (s > 0) {
                const svh_stage_t &P = stages[s0 + s - 1];
                const int tp = lds_swz(thr_local(P, tid));
                if (s > 1) __syncthreads();
#pragma unroll
                for (int j = 0; j < NR; ++j) lds[tp ^ P.lsw[j]] = v[j];
                __syncthreads();
                const int tn = lds_swz(thr_local(S, tid));
#pragma unroll
                for (int j = 0; j < NR; ++j) v[j] = lds[tn ^ S.lsw[j]];
            }
            for (int o = S.op_begin; o < S.op_end; ++o) {
                double g = 0.0;
                bool has_g = false;
                apply_op<R, RB, false>(v, w, ops[o], trig, &g, &has_g);
            }
        }
        {
            const svh_stage_t &S = stages[s0 + n_stages - 1];
            cpx<R> *p = psi + (base | thr_global<TB>(S, tid));
#pragma unroll
            for (int j = 0; j < NR; ++j) p[S.goff[j]] = v[j];
        }
    }
}

// (cos θ_p, sin θ_p) for every parameter, evaluated in FP64 and rounded once.
template <typename R>
__global__ void k_trig(const double *__restrict__ theta, int64_t np, cpx<R> *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) {
        double sn, cs;
        sincos(theta[i], &sn, &cs);
        out[i] = cpx<R>{(R)cs, (R)sn};
    }
}

// Gradient partials of up to kGradChunk passes, reduced after the sweep: slot
// s of pass entry e has its nb rows at part + off[e] + row·ng + (s - slot0).
constexpr int kGradChunk = 128;
struct GradChunk {
    int64_t off[kGradChunk];
    int32_t nb[kGradChunk];
    int32_t ng[kGradChunk];
    int32_t slot0[kGradChunk];
    int32_t npass;
    int32_t pad;
};

// Σ_rows partials of slot s (fixed-order tree); grid = the chunk's slots,
// contiguous from chunk.slot0[0].  With grad != nullptr every slot has its own
// parameter and the sum is added to grad[slot_param[s]] directly; otherwise it
// goes to slotsum[s] for k_grad_scatter.
__global__ void __launch_bounds__(256)
k_grad_slots(const double *__restrict__ part, GradChunk C, double *__restrict__ slotsum,
             const int32_t *__restrict__ slot_param, double *__restrict__ grad) {
    const int s = C.slot0[0] + (int)blockIdx.x;
    int e = 0;
    while (e + 1 < C.npass && C.slot0[e + 1] <= s) ++e;
    const int ng = C.ng[e], l = s - C.slot0[e];
    const double *p = part + C.off[e];
    double acc = 0.0;
    for (int b = threadIdx.x; b < C.nb[e]; b += blockDim.x) acc += p[(uint64_t)b * ng + l];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (grad) grad[slot_param[s]] += red[0];
        else slotsum[s] = red[0];
    }
}

// grad[param(s)] += slotsum[s] for s ascending (one thread: a fixed order even
// when several gates share a parameter).
__global__ void k_grad_scatter(const double *__restrict__ slotsum, int nslots,
                               const int32_t *__restrict__ slot_param, double *__restrict__ grad) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int s = 0; s < nslots; ++s) grad[slot_param[s]] += slotsum[s];
}


// ------------------------------------------------ host launch
constexpr size_t kMaxLds = 160 * 1024;

template <typename R, int TB>
static int launch_tb_pf(cpx<R> *psi, int n, const svh_pass_t &hp, const svh_pass_t *dp, const svh_stage_t *st,
                        const svh_op_t *op, const cpx<R> *trig, hipStream_t stream) {
    constexpr int L = TB + 4;
    const size_t lds = 2 * sizeof(cpx<R>) * ((size_t)1 << L);
    if (lds > kMaxLds) return fail(-7, "prefetch tile pass needs " + std::to_string(lds) + " B of LDS");
    auto kern = k_tile_pf<R, TB>;
    if (int e = ensure_lds_limit((const void *)kern, (int)kMaxLds, "hipFuncSetAttribute(k_tile_pf)")) return e;
    int resident = 0, cus = 0;  // workgroups per CU (occupancy calculator) and CUs of this device
    if (int e = resident_shape((const void *)kern, 1 << TB, lds, &resident, &cus)) return e;
    const uint64_t ntiles = 1ull << (n - L);
    const uint64_t cap = (uint64_t)resident * (uint64_t)cus;
    const uint64_t grid = ntiles < cap ? ntiles : cap;
    kern<<<(unsigned)grid, 1 << TB, lds, stream>>>(psi, n, dp, st, op, trig, ntiles);
    return checked_launch("k_tile_pf");
}

template <typename R, int TB, bool DUAL>
static int launch_tb(cpx<R> *psi, cpx<R> *lam, int n, const svh_pass_t &hp, const svh_pass_t *dp,
                     const svh_stage_t *st, const svh_op_t *op, const cpx<R> *trig, double *gpart,
                     hipStream_t stream) {
    constexpr int RB = 4;
    constexpr int L = TB + RB;
    constexpr int NW = ((1 << TB) + 63) / 64;
    const size_t lds = (DUAL ? 2 : 1) * sizeof(cpx<R>) * ((size_t)1 << L) +
                       (DUAL ? sizeof(double) * NW * (size_t)hp.n_grad : 0);
    if (lds > kMaxLds) return fail(-7, "tile pass needs " + std::to_string(lds) + " B of LDS (> 160 KiB)");
    auto kern = k_tile<R, TB, RB, DUAL>;
    if (int e = ensure_lds_limit((const void *)kern, (int)kMaxLds, "hipFuncSetAttribute(k_tile)")) return e;
    const uint64_t grid = 1ull << (n - L);
    if (grid > 0x7fffffffull) return fail(-7, "grid too large");
    kern<<<(unsigned)grid, 1 << TB, lds, stream>>>(psi, lam, n, dp, st, op, trig, gpart);
    return checked_launch("k_tile");
}

// Instantiated (TB, dtype, DUAL) combinations: the ones that fit in registers
// without spilling (c64: TB <= 10 forward, <= 9 adjoint; c128: TB <= 9 forward,
// <= 8 adjoint).
template <typename R, bool DUAL>
static int launch_tile(cpx<R> *psi, cpx<R> *lam, int n, const svh_pass_t &hp, const svh_pass_t *dp,
                       const svh_stage_t *st, const svh_op_t *op, const cpx<R> *trig, double *gpart,
                       hipStream_t stream) {
    if (hp.rb != 4) return fail(-7, "only 4 register bits are built");
    constexpr bool C64 = std::is_same<R, float>::value;
    constexpr int tb_max = DUAL ? (C64 ? 9 : 8) : (C64 ? 10 : 9);
    constexpr int tb_min = 4;
    if (hp.tb < tb_min || hp.tb > tb_max)
        return fail(-7, "tile pass tb=" + std::to_string(hp.tb) + " outside [" + std::to_string(tb_min) + ", " +
                            std::to_string(tb_max) + "] for this dtype/mode");
    if constexpr (!DUAL) {
        if (hp.flags & SVH_PASSF_PREFETCH) {
            if (hp.bits[0] != 0) return fail(-7, "prefetch passes need address bit 0 in the tile");
            switch (hp.tb) {
                case 6: return launch_tb_pf<R, 6>(psi, n, hp, dp, st, op, trig, stream);
                case 7: return launch_tb_pf<R, 7>(psi, n, hp, dp, st, op, trig, stream);
                case 8: return launch_tb_pf<R, 8>(psi, n, hp, dp, st, op, trig, stream);
                case 9: if constexpr (C64) return launch_tb_pf<R, 9>(psi, n, hp, dp, st, op, trig, stream); break;
                default: break;
            }
            return fail(-7, "prefetch tile pass tb must be in [6, 8] (c128) or [6, 9] (c64)");
        }
    }
    switch (hp.tb) {
        case 4: return launch_tb<R, 4, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream);
        case 5: return launch_tb<R, 5, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream);
        case 6: return launch_tb<R, 6, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream);
        case 7: return launch_tb<R, 7, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream);
        case 8: return launch_tb<R, 8, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream);
        case 9: if constexpr (tb_max >= 9) return launch_tb<R, 9, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream); break;
        case 10: if constexpr (tb_max >= 10) return launch_tb<R, 10, DUAL>(psi, lam, n, hp, dp, st, op, trig, gpart, stream); break;
        default: break;
    }
    return fail(-7, "tile pass tb not built");
}

bool gtile_pattern_matches(int rb, bool dual, int pid, const svh_op_t *ops, int nops);

// Validate a program's host copy against n; returns 0 or an error.  mode: the
// direction the program is about to run in (0 forward, 1 adjoint sweep), which
// fixes the pattern library its generic stages refer to; -1 skips that check.
static int check_program(const void *prog_host, int n, const svh_prog_header_t **hdr, int mode) {
    if (!prog_host) return fail(-1, "program pointer is NULL");
    const auto *h = static_cast<const svh_prog_header_t *>(prog_host);
    if (h->magic != SVH_PROG_MAGIC) return fail(-8, "bad program magic");
    if (h->version != SVH_ABI_VERSION) return fail(-8, "program version mismatch");
    if (n >= 0 && h->n_qubits != n) return fail(-8, "program built for a different n");
    const char *b = static_cast<const char *>(prog_host);
    const auto *passes = reinterpret_cast<const svh_pass_t *>(b + h->off_passes);
    const auto *stages = reinterpret_cast<const svh_stage_t *>(b + h->off_stages);
    const auto *ops = reinterpret_cast<const svh_op_t *>(b + h->off_ops);
    for (int p = 0; p < h->n_passes; ++p) {
        const svh_pass_t &P = passes[p];
        if (P.kind == SVH_PASS_TILE) {
            if (P.rb < 3 || P.rb > SVH_MAX_REG_BITS)
                return fail(-8, "tile pass " + std::to_string(p) + ": register bits must be 3, 4 or 5");
            if (P.L != P.tb + P.rb || P.L > h->n_qubits || P.L > SVH_MAX_TILE_BITS)
                return fail(-8, "tile pass " + std::to_string(p) + ": bad L");
            const bool generic = (P.flags & SVH_PASSF_GENERIC) != 0;
            if (generic && (P.rb < 3 || P.rb > 4 || P.tb < 6 || P.tb > 9 ||
                            (P.flags & (SVH_PASSF_HALF_LDS | SVH_PASSF_QUARTER_LDS | SVH_PASSF_PREFETCH))))
                return fail(-8, "tile pass " + std::to_string(p) +
                                    ": generic passes need rb 3 or 4, 6 <= tb <= 9, no LDS split or prefetch");
            if ((P.flags & SVH_PASSF_GEN_ZERO) && (p != 0 || mode == 1 || generic))
                return fail(-8, "tile pass " + std::to_string(p) +
                                    ": only the first pass of a forward program may generate |0...0>");
            const int extra = (P.flags & SVH_PASSF_REMAP_STORE) ? 1 : 0;  // the store record
            if (P.n_stages < 1 || P.stage_begin < 0 || P.stage_begin + P.n_stages + extra > h->n_stages)
                return fail(-8, "tile pass " + std::to_string(p) + ": bad stage range");
            if (extra) {
                const svh_stage_t &A = stages[P.stage_begin + P.n_stages - 1];
                const svh_stage_t &B = stages[P.stage_begin + P.n_stages];
                for (int k = 0; k < P.rb; ++k)
                    if (A.rbits[k] != B.rbits[k])
                        return fail(-8, "tile pass " + std::to_string(p) + ": store record layout differs");
                if (B.op_begin != B.op_end) return fail(-8, "tile pass " + std::to_string(p) + ": store record has ops");
            }
            uint64_t m = 0;
            for (int i = 0; i < P.L; ++i) {
                if (P.bits[i] < 0 || P.bits[i] >= h->n_qubits || ((m >> P.bits[i]) & 1))
                    return fail(-8, "tile pass " + std::to_string(p) + ": bad tile bits");
                m |= 1ull << P.bits[i];
            }
            for (int s = P.stage_begin; s < P.stage_begin + P.n_stages + extra; ++s) {
                const svh_stage_t &S = stages[s];
                int rm = 0;
                for (int k = 0; k < P.rb; ++k) {
                    if (S.rbits[k] < 0 || S.rbits[k] >= P.L || ((rm >> S.rbits[k]) & 1))
                        return fail(-8, "stage " + std::to_string(s) + ": bad register bits");
                    rm |= 1 << S.rbits[k];
                }
                if (S.op_begin < 0 || S.op_end < S.op_begin || S.op_end > h->n_ops)
                    return fail(-8, "stage " + std::to_string(s) + ": bad op range");
                // the precomputed address tables must describe this layout exactly
                int rs_prev = -1;
                for (int k = 0; k < P.rb; ++k) {
                    if (S.rsorted[k] <= rs_prev || !((rm >> S.rsorted[k]) & 1))
                        return fail(-8, "stage " + std::to_string(s) + ": bad sorted register bits");
                    rs_prev = S.rsorted[k];
                }
                uint64_t seen = 0;
                for (int i = 0; i < P.tb; ++i) {
                    const int g = S.tgbit[i];
                    if (g < 0 || g >= h->n_qubits || !((m >> g) & 1) || ((seen >> g) & 1))
                        return fail(-8, "stage " + std::to_string(s) + ": bad thread bit table");
                    seen |= 1ull << g;
                }
                for (int j = 0; j < (1 << P.rb); ++j) {
                    const uint64_t go = (uint64_t)S.goff[j];
                    if (S.goff[j] < 0 || (go & ~m) || (go & seen))
                        return fail(-8, "stage " + std::to_string(s) + ": bad register offset table");
                    if (S.lsw[j] < 0 || S.lsw[j] >= (1 << P.L))
                        return fail(-8, "stage " + std::to_string(s) + ": bad LDS offset table");
                }
                if (generic && S.pattern >= 0 && mode >= 0 &&
                    (s >= P.stage_begin + P.n_stages ||
                     !gtile_pattern_matches(P.rb, mode == 1, S.pattern, ops + S.op_begin, S.op_end - S.op_begin)))
                    return fail(-8, "stage " + std::to_string(s) + ": pattern id does not match its op sequence");
                int last_slot = -1;  // gradient terms of a stage are stored as one group
                for (int o = S.op_begin; o < S.op_end; ++o) {
                    const svh_op_t &O = ops[o];
                    if (O.flags & SVH_OPF_GRAD) {
                        if (last_slot >= 0 && O.grad_slot != last_slot + 1)
                            return fail(-8, "op " + std::to_string(o) +
                                                ": the gradient slots of a stage must be consecutive");
                        last_slot = O.grad_slot;
                    }
                    const bool two = O.kind == SVH_OP_MAT2 || O.kind == SVH_OP_CX ||
                                     O.kind == SVH_OP_CZ || O.kind == SVH_OP_ROT2;
                    if (O.kind < SVH_OP_MAT1 || O.kind > SVH_OP_ROT2)
                        return fail(-8, "op " + std::to_string(o) + ": bad kind");
                    if (O.a < 0 || O.a >= P.rb || (two && (O.b < 0 || O.b >= P.rb || O.b == O.a)))
                        return fail(-8, "op " + std::to_string(o) + ": bad register bit");
                    if ((O.kind == SVH_OP_MAT2 || O.kind == SVH_OP_ROT2) && O.a < O.b)
                        return fail(-8, "op " + std::to_string(o) + ": a must exceed b");
                    if ((O.kind == SVH_OP_ROT1 || O.kind == SVH_OP_ROT2) &&
                        (O.pauli < 1 || O.pauli > 3 || O.param < 0 || O.param >= h->n_params))
                        return fail(-8, "op " + std::to_string(o) + ": bad pauli or param index");
                    if ((O.flags & SVH_OPF_GRAD) && (O.grad_slot < 0 || O.grad_slot >= P.n_grad))
                        return fail(-8, "op " + std::to_string(o) + ": bad grad slot");
                }
            }
        } else if (P.kind == SVH_PASS_DENSE) {
            if (P.L < 1 || P.L > 6 || P.L > h->n_qubits) return fail(-8, "dense pass: bad k");
            if (P.pool_off < 0 || P.pool_off + 2 * (1ll << (2 * P.L)) > h->n_pool)
                return fail(-8, "dense pass: matrix outside pool");
        } else if (P.kind != SVH_PASS_NORMALIZE) {
            return fail(-8, "bad pass kind");
        }
    }
    if (h->off_perm) {
        const int32_t *perm = reinterpret_cast<const int32_t *>(b + h->off_perm);
        uint64_t seen = 0;
        for (int q = 0; q < h->n_qubits; ++q) {
            if (perm[q] < 0 || perm[q] >= h->n_qubits || ((seen >> perm[q]) & 1))
                return fail(-8, "final qubit permutation is not a permutation");
            seen |= 1ull << perm[q];
        }
    }
    if (h->off_cx) {
        if (!h->off_perm || mode == 1) return fail(-8, "final CNOTs need a final permutation (forward programs only)");
        const int32_t *cx = reinterpret_cast<const int32_t *>(b + h->off_cx);
        if (cx[0] < 1 || cx[0] > kMaxFinalCx) return fail(-8, "bad final CNOT count");
        for (int k = 0; k < cx[0]; ++k) {
            const int c = cx[1 + 2 * k], t = cx[2 + 2 * k];
            if (c < 0 || c >= h->n_qubits || t < 0 || t >= h->n_qubits || c == t)
                return fail(-8, "bad final CNOT bits");
        }
    }
    *hdr = h;
    return 0;
}

// ------------------------------------------------ bit permutation (out of place)
// dst[i'] = src[i] where bit b of i becomes bit perm[b] of i'.  A workgroup
// moves a tile of 2^TB amplitudes through LDS: the tile bits are the source
// bits 0..3, the source bits that land on destination bits 0..3, and the
// lowest other source bits, so both the loads and the stores run along 16
// consecutive amplitudes.
struct PermArgs {
    int32_t tsrc[kPermTileBits];  // tile-local bit x -> source bit (ascending)
    int32_t uidx[kPermTileBits];  // store-order bit i -> tile-local bit
    int32_t udst[kPermTileBits];  // store-order bit i -> destination bit (ascending)
    int32_t nsrc[64];             // non-tile source bits (ascending)
    int32_t ndst[64];             // their destination bits
    int32_t nnon;
    int32_t tbits;
    // CNOTs applied before the permutation, in gate order: the source of
    // destination element j is G1(G2(...Gk(j))) with Gi: a ^= bit(a, cxc[i]) << cxt[i]
    int32_t ncx;
    int32_t cxc[kMaxFinalCx];
    int32_t cxt[kMaxFinalCx];
};

template <typename R>
__global__ void __launch_bounds__(256) k_permute_bits(const cpx<R> *__restrict__ src, cpx<R> *__restrict__ dst,
                                                      PermArgs A) {
    __shared__ cpx<R> tile[1 << kPermTileBits];
    const int tid = threadIdx.x;
    const int TB = A.tbits;
    uint64_t sb = 0, db = 0, t = blockIdx.x;
    for (int k = 0; k < A.nnon && t; ++k, t >>= 1) {
        sb |= (t & 1ull) << A.nsrc[k];
        db |= (t & 1ull) << A.ndst[k];
    }
    // Element e = r·256 + tid (TB >= 8).  Every offset below is linear over
    // GF(2) in e and the bits of tid and r·256 are disjoint, so the thread's
    // part is computed once and the workgroup-uniform part once per r.
    const int per = (1 << TB) / 256;
    uint64_t so_t = 0, do_t = 0;
    int x_t = 0;
    for (int b = 0; b < 8; ++b) {
        const int bit = (tid >> b) & 1;
        so_t |= (uint64_t)bit << A.tsrc[b];
        x_t |= bit << A.uidx[b];
        do_t |= (uint64_t)bit << A.udst[b];
    }
    for (int r = 0; r < per; ++r) {
        uint64_t so = so_t;
        for (int b = 8; b < TB; ++b) so |= (uint64_t)((r >> (b - 8)) & 1) << A.tsrc[b];
        const int e = r * 256 + tid;
        uint64_t a = sb | so;
        for (int k = A.ncx - 1; k >= 0; --k) a ^= ((a >> A.cxc[k]) & 1ull) << A.cxt[k];
        tile[e ^ ((e >> 5) & 31)] = src[a];
    }
    __syncthreads();
    for (int r = 0; r < per; ++r) {
        int x = x_t;
        uint64_t dof = do_t;
        for (int b = 8; b < TB; ++b) {
            const int bit = (r >> (b - 8)) & 1;
            x |= bit << A.uidx[b];
            dof |= (uint64_t)bit << A.udst[b];
        }
        dst[db | dof] = tile[x ^ ((x >> 5) & 31)];
    }
}

static int permute_args(int n, const int *perm, PermArgs *A, int ncx = 0, const int32_t *cx = nullptr) {
    const int TB = n < kPermTileBits ? n : kPermTileBits;
    if (TB < 8) return fail(-2, "bit permutation needs n >= 8");
    std::vector<int> inv(n, -1);
    for (int b = 0; b < n; ++b) {
        if (perm[b] < 0 || perm[b] >= n || inv[perm[b]] >= 0) return fail(-4, "perm is not a permutation");
        inv[perm[b]] = b;
    }
    uint64_t tm = 0;
    int cnt = 0;
    auto add = [&](int b) {
        if (!((tm >> b) & 1) && cnt < TB) {
            tm |= 1ull << b;
            ++cnt;
        }
    };
    if (ncx < 0 || ncx > kMaxFinalCx) return fail(-4, "too many final CNOTs");
    for (int k = 0; k < ncx; ++k) {
        const int c = cx[2 * k], t = cx[2 * k + 1];
        if (c < 0 || c >= n || t < 0 || t >= n || c == t) return fail(-4, "bad final CNOT bits");
        A->cxc[k] = c;
        A->cxt[k] = t;
    }
    A->ncx = ncx;
    for (int b = 0; b < 4; ++b) add(b);
    for (int d = 0; d < 4; ++d) add(inv[d]);
    // CNOT targets inside the tile keep every tile's source addresses within
    // that tile (a CNOT only flips its target bit), so the gather stays coalesced
    for (int k = 0; k < ncx; ++k) add(cx[2 * k + 1]);
    for (int b = 0; b < n && cnt < TB; ++b) add(b);
    int x = 0;
    for (int b = 0; b < n; ++b)
        if ((tm >> b) & 1) A->tsrc[x++] = b;
    // store order: tile bits by destination bit
    int order[kPermTileBits];
    for (int i = 0; i < TB; ++i) order[i] = i;
    std::sort(order, order + TB, [&](int a, int c) { return perm[A->tsrc[a]] < perm[A->tsrc[c]]; });
    for (int i = 0; i < TB; ++i) {
        A->uidx[i] = order[i];
        A->udst[i] = perm[A->tsrc[order[i]]];
    }
    A->nnon = 0;
    for (int b = 0; b < n; ++b)
        if (!((tm >> b) & 1)) {
            A->nsrc[A->nnon] = b;
            A->ndst[A->nnon] = perm[b];
            A->nnon++;
        }
    A->tbits = TB;
    return 0;
}

static int launch_permute(const void *src, void *dst, int n, int dtype, const int *perm, hipStream_t st,
                          int ncx = 0, const int32_t *cx = nullptr) {
    PermArgs A;
    if (int e = permute_args(n, perm, &A, ncx, cx)) return e;
    const unsigned grid = (unsigned)(1ull << (n - A.tbits));
    if (dtype == SVH_C64)
        k_permute_bits<float><<<grid, 256, 0, st>>>((const cpx<float> *)src, (cpx<float> *)dst, A);
    else
        k_permute_bits<double><<<grid, 256, 0, st>>>((const cpx<double> *)src, (cpx<double> *)dst, A);
    return checked_launch("k_permute_bits");
}

// Dense passes take their matrix from the (double) pool; c64 needs a float copy.
template <typename R>
__global__ void k_pool_to(const double *__restrict__ src, cpx<R> *__restrict__ dst, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = cpx<R>{(R)src[2 * i], (R)src[2 * i + 1]};
}

}  // namespace svh

using namespace svh;

extern "C" int svh_apply_dense(void *psi, int n, int dtype, const int *bits, int k, const void *mat_dev,
                               void *stream);
extern "C" int svh_norm2(const void *psi, int n, int dtype, double *out_dev, void *scratch,
                         size_t scratch_bytes, void *stream);
extern "C" int svh_scale_by_inv_norm(void *psi, int n, int dtype, const double *norm2_dev, void *stream);
extern "C" int svh_apply_block5(void *psi, int n, int dtype, const int *bits, const void *u_dev, void *stream);
extern "C" int svh_apply_block5x2(void *psi, int n, int dtype, const int *bits_a, const void *ua_dev,
                                  const int *bits_b, const void *ub_dev, void *stream);

static bool dense_bits_disjoint(const svh_pass_t &a, const svh_pass_t &b) {
    uint64_t ma = 0, mb = 0;
    for (int j = 0; j < a.L; ++j) ma |= 1ull << a.bits[j];
    for (int j = 0; j < b.L; ++j) mb |= 1ull << b.bits[j];
    return (ma & mb) == 0;
}

// A run of consecutive complex64 5-qubit dense passes on pairwise disjoint
// bits (so they commute) on the matrix cores.  They are ordered by their
// highest bit and the k-th highest is paired with the k-th lowest: a pair's
// tile then takes its two spare bits from the low end of the address, which
// keeps the tile's HBM runs long and its span over pages small (pairing the
// two highest blocks would spread each tile over 2^10 distant pages).  An odd
// one out runs alone.  mat = scratch for two 32x32 complex64 matrices.
static int run_dense5_group(const svh_pass_t *g, int count, void *psi, int n, const double *dpool, void *mat,
                            hipStream_t st) {
    int order[8], top[8];
    for (int i = 0; i < count; ++i) {
        order[i] = i;
        top[i] = 0;
        for (int j = 0; j < 5; ++j) top[i] = g[i].bits[j] > top[i] ? g[i].bits[j] : top[i];
    }
    std::sort(order, order + count, [&](int a, int b) { return top[a] > top[b]; });
    void *mat2 = static_cast<char *>(mat) + 8 * 1024;
    for (int lo = 0, hi = count - 1; lo <= hi; ++lo, --hi) {
        const svh_pass_t &A = g[order[lo]];
        int ab[5];
        for (int j = 0; j < 5; ++j) ab[j] = A.bits[4 - j];  // row bit j = matrix qubit 4-j
        k_pool_to<float><<<4, 256, 0, st>>>(dpool + A.pool_off, (cpx<float> *)mat, 1024);
        if (int e = checked_launch("k_pool_to")) return e;
        int e = 0;
        if (lo == hi) {
            e = svh_apply_block5(psi, n, SVH_C64, ab, mat, st);
        } else {
            const svh_pass_t &B = g[order[hi]];
            int bb[5];
            for (int j = 0; j < 5; ++j) bb[j] = B.bits[4 - j];
            k_pool_to<float><<<4, 256, 0, st>>>(dpool + B.pool_off, (cpx<float> *)mat2, 1024);
            if ((e = checked_launch("k_pool_to"))) return e;
            e = svh_apply_block5x2(psi, n, SVH_C64, ab, mat, bb, mat2, st);
        }
        if (e) return e;
    }
    return 0;
}

// Scratch layout: [trig table: n_params (cos, sin) pairs][work area], the work
// area holding FP64 reduction partials, a dense pass's matrix, or the adjoint's
// per-workgroup gradient partials.
static size_t trig_bytes(const svh_prog_header_t *h, int dtype) {
    const size_t es = dtype == SVH_C64 ? 8 : 16;
    return ((size_t)h->n_params * es + 255) & ~(size_t)255;
}

namespace svh {
int gtile_launch(const void *psi, void *out, void *lam, int n, int dtype, const svh_pass_t &hp,
                 const svh_pass_t *dp, const svh_stage_t *dst, const svh_op_t *dop, const void *trig,
                 double *gpart, hipStream_t stream);
}  // namespace svh

// Rows of gradient partials an adjoint tile pass writes: one per tile, or one
// per wave in the generic kernels.
static uint64_t grad_rows(const svh_pass_t &P, int n) {
    const uint64_t tiles = 1ull << (n - P.L);
    return (P.flags & SVH_PASSF_GENERIC) ? tiles << (P.tb - 6) : tiles;
}

// Doubles of gradient partials of all tile passes of an adjoint program (each
// pass has its own region: rows·ng), plus one sum per gradient slot.
static size_t grad_doubles(const svh_prog_header_t *h) {
    const char *b = reinterpret_cast<const char *>(h);
    const auto *passes = reinterpret_cast<const svh_pass_t *>(b + h->off_passes);
    size_t d = (size_t)h->n_grad_slots;
    for (int p = 0; p < h->n_passes; ++p) {
        const svh_pass_t &P = passes[p];
        if (P.kind == SVH_PASS_TILE && P.n_grad > 0) d += grad_rows(P, h->n_qubits) * P.n_grad;
    }
    return d;
}

static size_t scratch_need(const svh_prog_header_t *h, int dtype, bool dual) {
    const char *b = reinterpret_cast<const char *>(h);
    const auto *passes = reinterpret_cast<const svh_pass_t *>(b + h->off_passes);
    size_t need = (kReduceBlocks + 8) * sizeof(double);  // norm partials + result
    const size_t es = dtype == SVH_C64 ? 8 : 16;
    for (int p = 0; p < h->n_passes; ++p) {
        const svh_pass_t &P = passes[p];
        size_t s = (kReduceBlocks + 8) * sizeof(double);
        if (P.kind == SVH_PASS_DENSE) s += 2 * es * ((size_t)1 << (2 * P.L));  // room for a pair
        if (s > need) need = s;
    }
    if (dual) {
        const size_t g = sizeof(double) * grad_doubles(h);
        if (g > need) need = g;
    }
    // a remapped program ends with a permutation through one state-sized buffer
    const size_t perm = h->off_perm ? es * ((size_t)1 << h->n_qubits) : 0;
    return trig_bytes(h, dtype) + ((need + 255) & ~(size_t)255) + perm;
}

// Kernels of a specialised program: svh_sp_<p> for the passes it covers.
struct SpHandle {
    std::vector<hipModule_t> modules;
    std::vector<hipFunction_t> fn;  // per pass, nullptr = not specialised
};

// Launch a specialised pass: svh_sp_<p>(psi, n, pass, st, ops, trig) or, for
// the adjoint, svh_spd_<p>(psi, lam, n, pass, st, ops, trig, gpart).
static int launch_sp(hipFunction_t f, void *psi, void *lam, int n, int dtype, const svh_pass_t &hp,
                     const svh_pass_t *dp, const svh_stage_t *dst, const svh_op_t *dop, const svh_stage_t *hst,
                     const void *trig, double *gpart, hipStream_t st, void *out = nullptr) {
    const int L = hp.L;
    const size_t es = dtype == SVH_C64 ? 8 : 16;
    const size_t nw = ((size_t)1 << hp.tb) >= 64 ? ((size_t)1 << hp.tb) / 64 : 1;
    const int split = (hp.flags & SVH_PASSF_HALF_LDS) ? 1 : (hp.flags & SVH_PASSF_QUARTER_LDS) ? 2 : 0;
    const size_t tile_amps = (size_t)1 << (L - split);
    const size_t lds = (lam ? 2 : 1) * es * tile_amps + (lam ? 8 * nw * (size_t)hp.n_grad : 0);
    if (lds > kMaxLds) return fail(-7, "specialised pass needs more than 160 KiB of LDS");
    const svh_stage_t *stp = dst + hp.stage_begin;
    const svh_op_t *opp = dop + hst[hp.stage_begin].op_begin;
    int nn = n;
    const unsigned grid = (unsigned)(1ull << (n - L));
    if (lam) {
        void *args[] = {&psi, &lam, &nn, (void *)&dp, (void *)&stp, (void *)&opp, (void *)&trig, (void *)&gpart};
        return hip_status(hipModuleLaunchKernel(f, grid, 1, 1, 1u << hp.tb, 1, 1, (unsigned)lds, st, args, nullptr),
                          "hipModuleLaunchKernel(svh_spd)");
    }
    if (out) {  // the out-of-place last pass of a remapped program
        void *args[] = {&psi, &out, &nn, (void *)&dp, (void *)&stp, (void *)&opp, (void *)&trig};
        return hip_status(hipModuleLaunchKernel(f, grid, 1, 1, 1u << hp.tb, 1, 1, (unsigned)lds, st, args, nullptr),
                          "hipModuleLaunchKernel(svh_sp)");
    }
    void *args[] = {&psi, &nn, (void *)&dp, (void *)&stp, (void *)&opp, (void *)&trig};
    return hip_status(hipModuleLaunchKernel(f, grid, 1, 1, 1u << hp.tb, 1, 1, (unsigned)lds, st, args, nullptr),
                      "hipModuleLaunchKernel(svh_sp)");
}

static int run_impl(void *psi, void *lam, int n, int dtype, const void *prog_host, const void *prog_dev,
                    const double *theta_dev, double *grad_dev, void *scratch, size_t scratch_bytes,
                    void *stream, bool dual, const SpHandle *sp = nullptr) {
    if (!psi || (dual && (!lam || !grad_dev))) return fail(-1, "state/lambda/grad pointer is NULL");
    if (n < 1 || n > kMaxQubits) return fail(-2, "n out of range");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "bad dtype");
    if (!prog_dev) return fail(-1, "device program pointer is NULL");
    const svh_prog_header_t *h = nullptr;
    if (int e = check_program(prog_host, n, &h, dual ? 1 : 0)) return e;
    const size_t need = scratch_need(h, dtype, dual);
    if (!scratch || scratch_bytes < need)
        return fail(-5, "scratch too small: need " + std::to_string(need) + " bytes");
    // a remapped program's final permutation goes through the scratch tail
    const size_t perm_bytes = h->off_perm ? (dtype == SVH_C64 ? 8 : 16) * ((size_t)1 << n) : 0;
    void *perm_tmp = h->off_perm ? static_cast<char *>(scratch) + (need - perm_bytes) : nullptr;
    bool last_in_tmp = false;
    if (h->n_params > 0 && !theta_dev) return fail(-1, "program has rotation ops but theta is NULL");
    const char *hb = static_cast<const char *>(prog_host);
    const char *db = static_cast<const char *>(prog_dev);
    const auto *hpass = reinterpret_cast<const svh_pass_t *>(hb + h->off_passes);
    const auto *dpass = reinterpret_cast<const svh_pass_t *>(db + h->off_passes);
    const auto *dstage = reinterpret_cast<const svh_stage_t *>(db + h->off_stages);
    const auto *dop = reinterpret_cast<const svh_op_t *>(db + h->off_ops);
    const auto *dgrad = reinterpret_cast<const int32_t *>(db + h->off_grad);
    const auto *dpool = reinterpret_cast<const double *>(db + h->off_pool);
    hipStream_t st = (hipStream_t)stream;
    void *trig = scratch;
    double *sd = reinterpret_cast<double *>(static_cast<char *>(scratch) + trig_bytes(h, dtype));

    if (h->n_params > 0) {
        const unsigned g = (unsigned)((h->n_params + 255) / 256);
        if (dtype == SVH_C64)
            k_trig<float><<<g, 256, 0, st>>>(theta_dev, h->n_params, (cpx<float> *)trig);
        else
            k_trig<double><<<g, 256, 0, st>>>(theta_dev, h->n_params, (cpx<double> *)trig);
        if (int e = checked_launch("k_trig")) return e;
    }

    const auto *hstage = reinterpret_cast<const svh_stage_t *>(hb + h->off_stages);
    if (sp && (int)sp->fn.size() != h->n_passes) return fail(-8, "specialised kernels do not match the program");
    // adjoint: every pass writes its gradient partials to its own region of sd;
    // they are reduced once after the sweep (two launches per chunk of passes)
    size_t goff = 0;
    GradChunk chunk{};
    double *slotsum = sd + (grad_doubles(h) - (size_t)h->n_grad_slots);
    bool distinct = true;  // every gradient slot has its own parameter
    if (dual) {
        const auto *hslot = reinterpret_cast<const int32_t *>(hb + h->off_grad);
        std::vector<char> seen((size_t)h->n_params, 0);
        for (int sl = 0; sl < h->n_grad_slots && distinct; ++sl) {
            const int q = hslot[sl];
            if (q < 0 || q >= h->n_params || seen[q]) distinct = false;
            else seen[q] = 1;
        }
    }
    auto flush_chunk = [&]() -> int {
        if (chunk.npass == 0) return 0;
        const int last = chunk.npass - 1;
        const int nslots = chunk.slot0[last] + chunk.ng[last] - chunk.slot0[0];
        k_grad_slots<<<nslots, 256, 0, st>>>(sd, chunk, slotsum, dgrad, distinct ? grad_dev : nullptr);
        chunk = GradChunk{};
        return checked_launch("k_grad_slots");
    };
    for (int p = 0; p < h->n_passes; ++p) {
        const svh_pass_t &P = hpass[p];
        int e = 0;
        double *gp = sd;
        if (dual && P.kind == SVH_PASS_TILE && P.n_grad > 0) {
            gp = sd + goff;
            if (chunk.npass == kGradChunk && (e = flush_chunk())) return e;
            chunk.off[chunk.npass] = (int64_t)goff;
            chunk.nb[chunk.npass] = (int32_t)grad_rows(P, n);
            chunk.ng[chunk.npass] = P.n_grad;
            chunk.slot0[chunk.npass] = P.grad_begin;
            chunk.npass++;
            goff += grad_rows(P, n) * P.n_grad;
        }
        if (P.kind == SVH_PASS_TILE && (P.flags & SVH_PASSF_GENERIC)) {
            // the generic kernels; a remapped program's last pass stores into
            // the scratch tail, which the final permutation reads
            void *out = (!dual && h->off_perm && p == h->n_passes - 1) ? perm_tmp : psi;
            e = gtile_launch(psi, out, dual ? lam : nullptr, n, dtype, P, dpass + p, dstage + P.stage_begin, dop,
                             trig, gp, st);
            if (out != psi) last_in_tmp = true;
        } else if ((P.flags & SVH_PASSF_GEN_ZERO) && !(P.kind == SVH_PASS_TILE && sp && sp->fn[p])) {
            // only the pass's compiled kernel generates |0...0>; any other
            // kernel would read a state the caller did not initialise
            return fail(-8, "pass 0 generates |0...0> and runs only in its compiled kernel (svh_run_program_sp)");
        } else if (P.kind == SVH_PASS_TILE && sp && sp->fn[p]) {
            // a remapped program's last compiled pass stores into the scratch
            // tail, which the final permutation reads
            void *out = (!dual && h->off_perm && p == h->n_passes - 1) ? perm_tmp : nullptr;
            e = launch_sp(sp->fn[p], psi, dual ? lam : nullptr, n, dtype, P, dpass + p, dstage, dop, hstage, trig,
                          gp, st, out);
            if (out) last_in_tmp = true;
        } else if (P.kind == SVH_PASS_TILE && (P.flags & SVH_PASSF_REMAP_STORE)) {
            return fail(-8, "remapped passes run only in their compiled kernels (svh_run_program_sp)");
        } else if (P.kind == SVH_PASS_TILE) {
            if (dtype == SVH_C64)
                e = dual ? launch_tile<float, true>((cpx<float> *)psi, (cpx<float> *)lam, n, P, dpass + p, dstage,
                                                    dop, (const cpx<float> *)trig, gp, st)
                         : launch_tile<float, false>((cpx<float> *)psi, nullptr, n, P, dpass + p, dstage, dop,
                                                     (const cpx<float> *)trig, sd, st);
            else
                e = dual ? launch_tile<double, true>((cpx<double> *)psi, (cpx<double> *)lam, n, P, dpass + p,
                                                     dstage, dop, (const cpx<double> *)trig, gp, st)
                         : launch_tile<double, false>((cpx<double> *)psi, nullptr, n, P, dpass + p, dstage, dop,
                                                      (const cpx<double> *)trig, sd, st);
        } else if (P.kind == SVH_PASS_DENSE && P.L == 5 && dtype == SVH_C64 && n >= 12) {
            if (dual) return fail(-8, "dense passes are not differentiable in the adjoint sweep");
            // complex64 5-qubit gates go to the matrix cores, two per sweep
            int q = p + 1;
            while (q < h->n_passes && hpass[q].kind == SVH_PASS_DENSE && hpass[q].L == 5 && q - p < 8) {
                bool ok = true;
                for (int r = p; r < q && ok; ++r) ok = dense_bits_disjoint(hpass[r], hpass[q]);
                if (!ok) break;
                ++q;
            }
            e = run_dense5_group(hpass + p, q - p, psi, n, dpool, sd + kReduceBlocks + 8, st);
            p = q - 1;
        } else if (P.kind == SVH_PASS_DENSE) {
            if (dual) return fail(-8, "dense passes are not differentiable in the adjoint sweep");
            const int K = 1 << P.L;
            void *mat = sd + kReduceBlocks + 8;
            if (dtype == SVH_C64)
                k_pool_to<float><<<(K * K + 255) / 256, 256, 0, st>>>(dpool + P.pool_off, (cpx<float> *)mat, K * K);
            else
                k_pool_to<double><<<(K * K + 255) / 256, 256, 0, st>>>(dpool + P.pool_off, (cpx<double> *)mat,
                                                                       K * K);
            e = checked_launch("k_pool_to");
            if (!e) e = svh_apply_dense(psi, n, dtype, P.bits, P.L, mat, stream);
        } else if (P.kind == SVH_PASS_NORMALIZE) {
            if (dual) return fail(-8, "normalisation is not differentiable in the adjoint sweep");
            e = svh_norm2(psi, n, dtype, sd + kReduceBlocks, sd, kReduceBlocks * sizeof(double), stream);
            if (!e) e = svh_scale_by_inv_norm(psi, n, dtype, sd + kReduceBlocks, stream);
        }
        if (e) return e;
    }
    if (dual && h->n_grad_slots > 0) {
        if (int e = flush_chunk()) return e;
        if (distinct) return 0;
        k_grad_scatter<<<1, 64, 0, st>>>(slotsum, h->n_grad_slots, dgrad, grad_dev);
        return checked_launch("k_grad_scatter");
    }
    if (h->off_perm) {
        // back to the natural order: physical bit perm[q] -> bit q.  The last
        // pass normally left the state in the scratch tail (one permutation
        // sweep into psi); otherwise permute into the tail and copy back.
        const int32_t *perm = reinterpret_cast<const int32_t *>(hb + h->off_perm);
        std::vector<int> back(n);
        for (int q = 0; q < n; ++q) back[perm[q]] = q;
        // CNOTs the program left to the permutation (count, then control/target pairs)
        const int32_t *cx = h->off_cx ? reinterpret_cast<const int32_t *>(hb + h->off_cx) : nullptr;
        const int ncx = cx ? cx[0] : 0;
        if (last_in_tmp) return launch_permute(perm_tmp, psi, n, dtype, back.data(), st, ncx, cx ? cx + 1 : nullptr);
        if (int e = launch_permute(psi, perm_tmp, n, dtype, back.data(), st, ncx, cx ? cx + 1 : nullptr)) return e;
        return hip_status(hipMemcpyAsync(psi, perm_tmp, perm_bytes, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
    }
    return 0;
}

template <typename R, int TB, bool DUAL>
static int launch_batch_tb(const svh_batch_item_t *items, int nitems, int64_t total_wg, int max_ngrad,
                           hipStream_t stream) {
    constexpr int L = TB + 4;
    constexpr int NW = ((1 << TB) + 63) / 64;
    const size_t lds = (DUAL ? 2 : 1) * sizeof(cpx<R>) * ((size_t)1 << L) +
                       (DUAL ? sizeof(double) * NW * (size_t)max_ngrad : 0);
    if (lds > kMaxLds) return fail(-7, "batched tile pass needs " + std::to_string(lds) + " B of LDS");
    auto kern = k_tile_batch<R, TB, 4, DUAL>;
    if (int e = ensure_lds_limit((const void *)kern, (int)kMaxLds, "hipFuncSetAttribute(k_tile_batch)")) return e;
    if (total_wg <= 0 || total_wg > 0x7fffffff) return fail(-4, "bad workgroup count");
    kern<<<(unsigned)total_wg, 1 << TB, lds, stream>>>(items, nitems);
    return checked_launch("k_tile_batch");
}

template <typename R, bool DUAL>
static int launch_batch(const svh_batch_item_t *items, int nitems, int64_t total_wg, int tb, int max_ngrad,
                        hipStream_t stream) {
    constexpr bool C64 = std::is_same<R, float>::value;
    constexpr int tb_max = DUAL ? (C64 ? 9 : 8) : (C64 ? 10 : 9);
    if (tb < 4 || tb > tb_max) return fail(-7, "batched tile pass: tb outside the built range");
    switch (tb) {
        case 4: return launch_batch_tb<R, 4, DUAL>(items, nitems, total_wg, max_ngrad, stream);
        case 5: return launch_batch_tb<R, 5, DUAL>(items, nitems, total_wg, max_ngrad, stream);
        case 6: return launch_batch_tb<R, 6, DUAL>(items, nitems, total_wg, max_ngrad, stream);
        case 7: return launch_batch_tb<R, 7, DUAL>(items, nitems, total_wg, max_ngrad, stream);
        case 8: return launch_batch_tb<R, 8, DUAL>(items, nitems, total_wg, max_ngrad, stream);
        case 9: if constexpr (tb_max >= 9) return launch_batch_tb<R, 9, DUAL>(items, nitems, total_wg, max_ngrad, stream); break;
        case 10: if constexpr (tb_max >= 10) return launch_batch_tb<R, 10, DUAL>(items, nitems, total_wg, max_ngrad, stream); break;
        default: break;
    }
    return fail(-7, "batched tile pass: tb not built");
}

extern "C" {

int svh_batch_tile(const void *items_dev, int nitems, int64_t total_wg, int dtype, int tb, int dual,
                   int max_ngrad, void *stream) {
    if (!items_dev || nitems <= 0) return fail(-1, "empty batch item table");
    if (max_ngrad < 0) return fail(-4, "bad max_ngrad");
    const auto *it = static_cast<const svh_batch_item_t *>(items_dev);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SVH_C64)
        return dual ? launch_batch<float, true>(it, nitems, total_wg, tb, max_ngrad, st)
                    : launch_batch<float, false>(it, nitems, total_wg, tb, max_ngrad, st);
    if (dtype == SVH_C128)
        return dual ? launch_batch<double, true>(it, nitems, total_wg, tb, max_ngrad, st)
                    : launch_batch<double, false>(it, nitems, total_wg, tb, max_ngrad, st);
    return fail(-3, "bad dtype");
}

int svh_program_scratch_bytes(const void *prog_host, int dtype, size_t *bytes) {
    if (!bytes) return fail(-1, "bytes pointer is NULL");
    const svh_prog_header_t *h = nullptr;
    if (int e = check_program(prog_host, -1, &h, -1)) return e;
    *bytes = scratch_need(h, dtype, true);
    return 0;
}

int svh_run_program(void *psi, int n, int dtype, const void *prog_host, const void *prog_dev,
                    const double *theta_dev, void *scratch, size_t scratch_bytes, void *stream) {
    return run_impl(psi, nullptr, n, dtype, prog_host, prog_dev, theta_dev, nullptr, scratch, scratch_bytes,
                    stream, false);
}

int svh_sp_load(const char *const *paths, int npaths, int n_passes, void **handle) {
    if (!paths || npaths <= 0 || n_passes < 0 || !handle) return fail(-1, "bad arguments");
    auto *H = new SpHandle();
    H->fn.assign(n_passes, nullptr);
    for (int i = 0; i < npaths; ++i) {
        hipModule_t m = nullptr;
        hipError_t e = hipModuleLoad(&m, paths[i]);
        if (e != hipSuccess) {
            for (auto mm : H->modules) (void)hipModuleUnload(mm);
            delete H;
            return hip_status(e, "hipModuleLoad");
        }
        H->modules.push_back(m);
        for (int p = 0; p < n_passes; ++p) {
            if (H->fn[p]) continue;
            hipFunction_t f = nullptr;
            const std::string name = "svh_sp_" + std::to_string(p);
            const std::string named = "svh_spd_" + std::to_string(p);
            if (hipModuleGetFunction(&f, m, name.c_str()) == hipSuccess) H->fn[p] = f;
            else if (hipModuleGetFunction(&f, m, named.c_str()) == hipSuccess) H->fn[p] = f;
        }
    }
    (void)hipGetLastError();  // lookups of absent kernels are expected
    *handle = H;
    return 0;
}

int svh_spb_load(const char *const *paths, int npaths, const char *const *names, int nnames, void **handle) {
    if (!paths || npaths <= 0 || !names || nnames <= 0 || !handle) return fail(-1, "bad arguments");
    auto *H = new SpHandle();
    H->fn.assign(nnames, nullptr);
    for (int i = 0; i < npaths; ++i) {
        hipModule_t m = nullptr;
        hipError_t e = hipModuleLoad(&m, paths[i]);
        if (e != hipSuccess) {
            for (auto mm : H->modules) (void)hipModuleUnload(mm);
            delete H;
            return hip_status(e, "hipModuleLoad");
        }
        H->modules.push_back(m);
        for (int k = 0; k < nnames; ++k) {
            if (H->fn[k] || !names[k]) continue;
            hipFunction_t f = nullptr;
            if (hipModuleGetFunction(&f, m, names[k]) == hipSuccess) H->fn[k] = f;
        }
    }
    (void)hipGetLastError();  // lookups of kernels in the other modules are expected to fail
    for (int k = 0; k < nnames; ++k) {
        if (!H->fn[k]) {
            const std::string nm = names[k] ? names[k] : "(null)";
            for (auto mm : H->modules) (void)hipModuleUnload(mm);
            delete H;
            return fail(-9, "batch kernel " + nm + " not found in the code objects");
        }
    }
    *handle = H;
    return 0;
}

int svh_spb_launch(void *handle, int index, const void *items_dev, int nitems, int64_t total_wg, int tb,
                   size_t lds_bytes, void *stream) {
    const auto *H = static_cast<const SpHandle *>(handle);
    if (!H || !items_dev || nitems <= 0) return fail(-1, "bad arguments");
    if (index < 0 || index >= (int)H->fn.size()) return fail(-4, "batch kernel index out of range");
    if (tb < 1 || tb > 10) return fail(-4, "bad workgroup size");
    if (total_wg <= 0 || total_wg > 0x7fffffff) return fail(-4, "bad workgroup count");
    if (lds_bytes > kMaxLds) return fail(-7, "batch kernel needs more than 160 KiB of LDS");
    const void *it = items_dev;
    int ni = nitems;
    void *args[] = {(void *)&it, &ni};
    return hip_status(hipModuleLaunchKernel(H->fn[index], (unsigned)total_wg, 1, 1, 1u << tb, 1, 1, (unsigned)lds_bytes,
                                            (hipStream_t)stream, args, nullptr),
                      "hipModuleLaunchKernel(svh_spb)");
}

int svh_sp_free(void *handle) {
    auto *H = static_cast<SpHandle *>(handle);
    if (!H) return 0;
    int rc = 0;
    for (auto m : H->modules) {
        const int e = hip_status(hipModuleUnload(m), "hipModuleUnload");
        if (e && !rc) rc = e;
    }
    delete H;
    return rc;
}

int svh_run_program_sp(void *handle, void *psi, int n, int dtype, const void *prog_host, const void *prog_dev,
                       const double *theta_dev, void *scratch, size_t scratch_bytes, void *stream) {
    if (!handle) return fail(-1, "specialised handle is NULL");
    return run_impl(psi, nullptr, n, dtype, prog_host, prog_dev, theta_dev, nullptr, scratch, scratch_bytes,
                    stream, false, static_cast<const SpHandle *>(handle));
}

int svh_run_adjoint_sp(void *handle, void *psi, void *lam, int n, int dtype, const void *prog_host,
                       const void *prog_dev, const double *theta_dev, double *grad_dev, void *scratch,
                       size_t scratch_bytes, void *stream) {
    if (!handle) return fail(-1, "specialised handle is NULL");
    return run_impl(psi, lam, n, dtype, prog_host, prog_dev, theta_dev, grad_dev, scratch, scratch_bytes, stream,
                    true, static_cast<const SpHandle *>(handle));
}

int svh_run_adjoint(void *psi, void *lam, int n, int dtype, const void *prog_host, const void *prog_dev,
                    const double *theta_dev, double *grad_dev, void *scratch, size_t scratch_bytes,
                    void *stream) {
    return run_impl(psi, lam, n, dtype, prog_host, prog_dev, theta_dev, grad_dev, scratch, scratch_bytes,
                    stream, true);
}

}  // extern "C"

// ------------------------------------------------ allocation-free program runner
// svh_run_program with the scratch buffer in a fixed device array instead of a
// caller buffer, for callers that cannot allocate (graph capture, serving
// loops).  Every device holds its own copy of the array for the life of the
// process; the host only ever uses its device address, looked up once per
// device (hipGetSymbolAddress) and kept here.
namespace svh {
constexpr size_t kStaticScratchBytes = 16u << 20;
}  // namespace svh
static __device__ __align__(256) unsigned char scratch[svh::kStaticScratchBytes];

static int static_scratch(void **out) {
    static std::mutex mu;
    static std::unordered_map<int, void *> ptrs;
    int dev = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) {
        *out = it->second;
        return 0;
    }
    void *p = nullptr;
    if (int e = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return e;
    ptrs[dev] = p;
    *out = p;
    return 0;
}

extern "C" int svh_run_program_static(void *psi, int n, int dtype, const void *prog_host, const void *prog_dev,
                                      const double *theta_dev, void *stream) {
    if (!psi) return fail(-1, "state pointer is NULL");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "bad dtype");
    const svh_prog_header_t *h = nullptr;
    if (int e = check_program(prog_host, n, &h, 0)) return e;
    const size_t need = scratch_need(h, dtype, false);
    if (need > kStaticScratchBytes)
        return fail(-5, "program needs " + std::to_string(need) + " bytes of scratch; the static scratch holds " +
                            std::to_string(kStaticScratchBytes));
    void *p = nullptr;
    if (int e = static_scratch(&p)) return e;
    return run_impl(psi, nullptr, n, dtype, prog_host, prog_dev, theta_dev, nullptr, p, kStaticScratchBytes, stream,
                    false);
}


extern "C" int svh_permute_bits(const void *src, void *dst, int n, int dtype, const int *perm, void *stream) {
    if (!src || !dst || !perm) return fail(-1, "NULL pointer");
    if (src == dst) return fail(-1, "the bit permutation is out of place: src and dst must differ");
    if (n < 8 || n > kMaxQubits) return fail(-2, "bit permutation needs 8 <= n <= 36");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "bad dtype");
    return launch_permute(src, dst, n, dtype, perm, (hipStream_t)stream);
}
