// Measurement-side kernels: marginal probabilities over a set of qubits,
// inverse-CDF sampling, the collapse gather, and the partial overlap
// M[i,j] = Σ_r a[i,r]·conj(b[j,r]) behind reduced density matrices and
// environment tensors.  Every reduction is a fixed-order FP64 tree, so results
// do not depend on scheduling.
//
// Index convention: a "kept" set is given as address bits, most significant
// first: kept[0] carries bit k-1 of the kept index m.  The remaining ("rest")
// address bits carry the rest index r in ascending order, so r is big-endian
// over the remaining qubits exactly like the reference's ncon output order.
#include "svh_common.h"

namespace svh {

struct KeptBits {
    int bit[kMaxQubits];
};

// Address bits of m (k bits, MSB first onto kept[0..k-1]).
__device__ __forceinline__ uint64_t deposit_kept(uint64_t m, const KeptBits &kb, int k) {
    uint64_t a = 0;
    for (int i = 0; i < k; ++i) a |= ((m >> (k - 1 - i)) & 1ull) << kb.bit[i];
    return a;
}

// Bits of r deposited into the set bits of mask, lowest first.
__device__ __forceinline__ uint64_t deposit_mask(uint64_t r, uint64_t mask) {
    uint64_t a = 0;
    while (mask && r) {
        const uint64_t low = mask & (~mask + 1);
        if (r & 1) a |= low;
        r >>= 1;
        mask &= mask - 1;
    }
    return a;
}

// x + d inside the bit subset mask (x, d ⊂ mask): the bits outside mask are
// set so that carries ripple across them.
__device__ __forceinline__ uint64_t masked_add(uint64_t x, uint64_t d, uint64_t mask) {
    return ((x | ~mask) + d) & mask;
}

template <typename R>
__device__ __forceinline__ double abs2(const cpx<R> &z) {
    return (double)z.x * z.x + (double)z.y * z.y;
}

// ---------------------------------------------------------------- marginals
// Few outcomes (M ≤ 4096): block (m, c) sums |ψ|² over the c-th chunk of the
// rest index with its 256 threads, then a tree; partial[c*M + m].
template <typename R>
__global__ void __launch_bounds__(256)
k_marg_block(const cpx<R> *__restrict__ psi, KeptBits kb, int k, uint64_t rest_mask, uint64_t M,
             uint64_t chunk, double *__restrict__ part) {
    const uint64_t m = blockIdx.x % M, c = blockIdx.x / M;
    const uint64_t base = deposit_kept(m, kb, k);
    const uint64_t r0 = c * chunk + threadIdx.x;
    double acc = 0.0;
    if (threadIdx.x < chunk) {
        uint64_t x = deposit_mask(r0, rest_mask);
        const uint64_t d = deposit_mask(256, rest_mask);
        for (uint64_t r = r0; r < (c + 1) * chunk; r += 256) {
            acc += abs2(psi[base | x]);
            x = masked_add(x, d, rest_mask);
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[c * M + m] = red[0];
}

// Many outcomes: thread (m, c) sums its chunk sequentially.
template <typename R>
__global__ void __launch_bounds__(256)
k_marg_thread(const cpx<R> *__restrict__ psi, KeptBits kb, int k, uint64_t rest_mask, uint64_t M,
              uint64_t C, uint64_t chunk, double *__restrict__ part) {
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= M * C) return;
    const uint64_t m = g % M, c = g / M;
    const uint64_t base = deposit_kept(m, kb, k);
    uint64_t x = deposit_mask(c * chunk, rest_mask);
    double acc = 0.0;
    for (uint64_t r = 0; r < chunk; ++r) {
        acc += abs2(psi[base | x]);
        x = masked_add(x, 1, rest_mask);
    }
    part[c * M + m] = acc;
}

// out[m] = Σ_c part[c*M + m], c ascending.
__global__ void __launch_bounds__(256) k_marg_final(const double *__restrict__ part, uint64_t M, uint64_t C,
                                                    double *__restrict__ out) {
    const uint64_t m = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    double s = 0.0;
    for (uint64_t c = 0; c < C; ++c) s += part[c * M + m];
    out[m] = s;
}

// ---------------------------------------------------------------- CDF + search
constexpr int kScanPer = 8;                  // elements per thread
constexpr int kScanTile = 256 * kScanPer;    // elements per block

// Exclusive scan of one value per thread over a 256-thread block (fixed
// Hillis-Steele order); returns this thread's exclusive prefix and the total.
__device__ __forceinline__ double block_exclusive_scan(double v, double &total) {
    __shared__ double buf[2][256];
    int cur = 0;
    buf[0][threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const double add = (int)threadIdx.x >= off ? buf[cur][threadIdx.x - off] : 0.0;
        buf[cur ^ 1][threadIdx.x] = buf[cur][threadIdx.x] + add;
        cur ^= 1;
        __syncthreads();
    }
    const double excl = threadIdx.x == 0 ? 0.0 : buf[cur][threadIdx.x - 1];
    total = buf[cur][255];
    __syncthreads();
    return excl;
}

__device__ __forceinline__ void load_tile(const double *__restrict__ p, uint64_t M, uint64_t i0,
                                          double (&v)[kScanPer], double &tsum) {
    tsum = 0.0;
#pragma unroll
    for (int e = 0; e < kScanPer; ++e) {
        const uint64_t i = i0 + e;
        v[e] = i < M ? p[i] : 0.0;
        tsum += v[e];
    }
}

__global__ void __launch_bounds__(256) k_scan_partial(const double *__restrict__ p, uint64_t M,
                                                      double *__restrict__ bsum) {
    double v[kScanPer], tsum, total;
    load_tile(p, M, (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer, v, tsum);
    block_exclusive_scan(tsum, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// One block: exclusive offsets of the nb block sums (each thread takes a
// contiguous run, sequentially).
__global__ void __launch_bounds__(256) k_scan_top(const double *__restrict__ bsum, uint64_t nb,
                                                  double *__restrict__ boff) {
    const uint64_t per = (nb + 255) / 256;
    const uint64_t b0 = threadIdx.x * per;
    double s = 0.0;
    for (uint64_t b = b0; b < b0 + per && b < nb; ++b) s += bsum[b];
    double total;
    double run = block_exclusive_scan(s, total);
    for (uint64_t b = b0; b < b0 + per && b < nb; ++b) {
        boff[b] = run;
        run += bsum[b];
    }
}

__global__ void __launch_bounds__(256) k_scan_final(const double *__restrict__ p, uint64_t M,
                                                    const double *__restrict__ boff, double *__restrict__ cdf) {
    double v[kScanPer], tsum, total;
    const uint64_t i0 = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    load_tile(p, M, i0, v, tsum);
    double run = boff[blockIdx.x] + block_exclusive_scan(tsum, total);
#pragma unroll
    for (int e = 0; e < kScanPer; ++e) {
        run += v[e];
        if (i0 + e < M) cdf[i0 + e] = run;
    }
}

// idx[s] = first i with cdf[i]/cdf[M-1] > u[s]  (numpy searchsorted, side='right',
// on the cdf normalised by its last element).
__global__ void __launch_bounds__(256) k_search(const double *__restrict__ cdf, uint64_t M,
                                                const double *__restrict__ u, uint64_t shots,
                                                int64_t *__restrict__ idx) {
    const uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= shots) return;
    const double last = cdf[M - 1];
    const double x = u[s];
    uint64_t lo = 0, hi = M;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (cdf[mid] / last > x) hi = mid;
        else lo = mid + 1;
    }
    idx[s] = (int64_t)(lo < M ? lo : M - 1);
}

__global__ void k_copy_last(const double *__restrict__ cdf, uint64_t M, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = cdf[M - 1];
}

// ---------------------------------------------------------------- collapse
template <typename R>
__global__ void __launch_bounds__(256)
k_gather_rest(const cpx<R> *__restrict__ psi, uint64_t base, uint64_t rest_mask, uint64_t NR,
              cpx<R> *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= NR) return;
    uint64_t x = deposit_mask(r, rest_mask);
    const uint64_t d = deposit_mask(stride, rest_mask);
    for (; r < NR; r += stride) {
        out[r] = psi[base | x];
        x = masked_add(x, d, rest_mask);
    }
}

// ---------------------------------------------------------------- partial overlap
// Stage 1: block g covers rest chunk g.  With P = M² ≤ 256 pairs, T = 256/P
// threads share a pair and stride over r; otherwise each thread owns P/256
// pairs.  part[g*P + p] holds (re, im) of the block's sum.
template <typename R, int Q>
__global__ void __launch_bounds__(256)
k_overlap(const cpx<R> *__restrict__ a, const cpx<R> *__restrict__ b, KeptBits kb, int k,
          uint64_t rest_mask, uint64_t chunk, double2 *__restrict__ part) {
    const int M = 1 << k, P = M * M;
    const int T = P >= 256 ? 1 : 256 / P;
    const int sub = P >= 256 ? 0 : (int)threadIdx.x / P;
    const uint64_t r_begin = (uint64_t)blockIdx.x * chunk;
    double2 acc[Q];
    uint64_t bi[Q], bj[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int p = P >= 256 ? (int)threadIdx.x + q * 256 : (int)threadIdx.x % P;
        acc[q] = make_double2(0.0, 0.0);
        bi[q] = deposit_kept((uint64_t)(p / M), kb, k);
        bj[q] = deposit_kept((uint64_t)(p % M), kb, k);
    }
    if (sub < (int)chunk) {
        uint64_t x = deposit_mask(r_begin + sub, rest_mask);
        const uint64_t d = deposit_mask((uint64_t)T, rest_mask);
        for (uint64_t r = sub; r < chunk; r += T) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const cpx<R> u = a[bi[q] | x], w = b[bj[q] | x];
                // u · conj(w)
                acc[q].x += (double)u.x * w.x + (double)u.y * w.y;
                acc[q].y += (double)u.y * w.x - (double)u.x * w.y;
            }
            x = masked_add(x, d, rest_mask);
        }
    }
    if (P >= 256) {
#pragma unroll
        for (int q = 0; q < Q; ++q) part[(uint64_t)blockIdx.x * P + threadIdx.x + q * 256] = acc[q];
        return;
    }
    // fixed-order tree over the T threads of each pair
    __shared__ double2 red[256];
    red[threadIdx.x] = acc[0];
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (sub < s) {
            const double2 o = red[threadIdx.x + s * P];
            red[threadIdx.x].x += o.x;
            red[threadIdx.x].y += o.y;
        }
        __syncthreads();
    }
    if (sub == 0) part[(uint64_t)blockIdx.x * P + threadIdx.x] = red[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_overlap_final(const double2 *__restrict__ part, int P, int G,
                                                       double2 *__restrict__ out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    double2 s = make_double2(0.0, 0.0);
    for (int g = 0; g < G; ++g) {
        const double2 v = part[(uint64_t)g * P + p];
        s.x += v.x;
        s.y += v.y;
    }
    out[p] = s;
}

// ---------------------------------------------------------------- host side
static int check_kept(int n, const int *kept, int k, KeptBits &kb, uint64_t &kmask) {
    if (n < 1 || n > kMaxQubits) return fail(-2, "n must be in [1, 36], got " + std::to_string(n));
    if (k < 0 || k > n) return fail(-2, "number of kept qubits must be in [0, n]");
    if (k > 0 && !kept) return fail(-1, "kept-bit array is NULL");
    kmask = 0;
    for (int i = 0; i < k; ++i) {
        const int b = kept[i];
        if (b < 0 || b >= n) return fail(-2, "kept bit " + std::to_string(b) + " outside the register");
        if ((kmask >> b) & 1) return fail(-2, "kept bit " + std::to_string(b) + " repeated");
        kmask |= 1ull << b;
        kb.bit[i] = b;
    }
    return 0;
}

struct MargPlan {
    bool block;
    uint64_t M, C, chunk;
};

static MargPlan marg_plan(int n, int k) {
    MargPlan p;
    p.M = 1ull << k;
    const uint64_t NR = 1ull << (n - k);
    if (p.M <= 4096) {
        p.block = true;
        uint64_t C = 4096 / p.M;
        const uint64_t cap = NR >= 1024 ? NR / 1024 : 1;
        if (C > cap) C = cap;
        p.C = C;
    } else {
        p.block = false;
        uint64_t C = p.M >= (1ull << 18) ? 1 : (1ull << 18) / p.M;
        if (C > NR) C = NR;
        p.C = C;
    }
    p.chunk = NR / p.C;
    return p;
}

static size_t scan_blocks(uint64_t M) { return (size_t)((M + kScanTile - 1) / kScanTile); }

static int overlap_blocks(int n, int k) {
    const uint64_t P = 1ull << (2 * k);
    const uint64_t NR = 1ull << (n - k);
    uint64_t G = (1ull << 18) / P;
    if (G > 1024) G = 1024;
    const uint64_t cap = NR >= 64 ? NR / 64 : 1;
    if (G > cap) G = cap;
    if (G < 1) G = 1;
    return (int)G;
}

}  // namespace svh

using namespace svh;

extern "C" {

size_t svh_measure_scratch_bytes(int what, int n, int k) {
    if (n < 1 || n > kMaxQubits || k < 0 || k > n) return 0;
    switch (what) {
    case SVH_MEAS_PROBS: {
        const MargPlan p = marg_plan(n, k);
        return p.C > 1 ? (size_t)(p.C * p.M * sizeof(double)) : sizeof(double);
    }
    case SVH_MEAS_SAMPLE: {
        const uint64_t M = 1ull << k;
        return (size_t)(M * sizeof(double) + 2 * scan_blocks(M) * sizeof(double));
    }
    case SVH_MEAS_OVERLAP:
        if (k > 6) return 0;
        return (size_t)overlap_blocks(n, k) * ((size_t)1 << (2 * k)) * sizeof(double2);
    default:
        return 0;
    }
}

int svh_marginal_probs(const void *psi, int n, int dtype, const int *kept, int k, double *probs,
                       void *scratch, size_t scratch_bytes, void *stream) {
    if (!psi || !probs) return fail(-1, "state or output pointer is NULL");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "dtype must be SVH_C64 or SVH_C128");
    KeptBits kb{};
    uint64_t kmask;
    if (int e = check_kept(n, kept, k, kb, kmask)) return e;
    const MargPlan p = marg_plan(n, k);
    const size_t need = svh_measure_scratch_bytes(SVH_MEAS_PROBS, n, k);
    if (p.C > 1 && (!scratch || scratch_bytes < need))
        return fail(-5, "scratch too small for svh_marginal_probs (need " + std::to_string(need) + " bytes)");
    const uint64_t rest_mask = ((1ull << n) - 1) & ~kmask;
    hipStream_t st = (hipStream_t)stream;
    double *part = p.C > 1 ? (double *)scratch : probs;
    if (p.block) {
        const unsigned g = (unsigned)(p.M * p.C);
        if (dtype == SVH_C64)
            k_marg_block<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, kb, k, rest_mask, p.M, p.chunk, part);
        else
            k_marg_block<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, kb, k, rest_mask, p.M, p.chunk, part);
    } else {
        const unsigned g = (unsigned)((p.M * p.C + 255) / 256);
        if (dtype == SVH_C64)
            k_marg_thread<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, kb, k, rest_mask, p.M, p.C, p.chunk,
                                                    part);
        else
            k_marg_thread<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, kb, k, rest_mask, p.M, p.C, p.chunk,
                                                     part);
    }
    if (p.C > 1) k_marg_final<<<(unsigned)((p.M + 255) / 256), 256, 0, st>>>(part, p.M, p.C, probs);
    return checked_launch("svh_marginal_probs");
}

int svh_sample_cdf(const double *probs, int k, const double *uniforms, int64_t shots, int64_t *idx,
                   double *total, void *scratch, size_t scratch_bytes, void *stream) {
    if (!probs || !total) return fail(-1, "probability or total pointer is NULL");
    if (k < 0 || k > kMaxQubits) return fail(-2, "k must be in [0, 36]");
    if (shots < 0) return fail(-4, "shots must be >= 0");
    if (shots > 0 && (!uniforms || !idx)) return fail(-1, "uniform or index pointer is NULL");
    const uint64_t M = 1ull << k;
    const size_t nb = scan_blocks(M);
    const size_t need = (size_t)(M * sizeof(double) + 2 * nb * sizeof(double));
    if (!scratch || scratch_bytes < need)
        return fail(-5, "scratch too small for svh_sample_cdf (need " + std::to_string(need) + " bytes)");
    if (nb > (1ull << 31)) return fail(-4, "too many outcomes");
    hipStream_t st = (hipStream_t)stream;
    double *cdf = (double *)scratch;
    double *bsum = cdf + M;
    double *boff = bsum + nb;
    k_scan_partial<<<(unsigned)nb, 256, 0, st>>>(probs, M, bsum);
    k_scan_top<<<1, 256, 0, st>>>(bsum, nb, boff);
    k_scan_final<<<(unsigned)nb, 256, 0, st>>>(probs, M, boff, cdf);
    k_copy_last<<<1, 64, 0, st>>>(cdf, M, total);
    if (shots > 0)
        k_search<<<(unsigned)((shots + 255) / 256), 256, 0, st>>>(cdf, M, uniforms, (uint64_t)shots, idx);
    return checked_launch("svh_sample_cdf");
}

int svh_collapse_gather(const void *psi, int n, int dtype, const int *kept, int k, uint64_t outcome, void *out,
                        void *stream) {
    if (!psi || !out) return fail(-1, "state or output pointer is NULL");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "dtype must be SVH_C64 or SVH_C128");
    KeptBits kb{};
    uint64_t kmask;
    if (int e = check_kept(n, kept, k, kb, kmask)) return e;
    if (k < 64 && (outcome >> k) != 0) return fail(-4, "outcome has bits beyond the measured qubits");
    uint64_t base = 0;
    for (int i = 0; i < k; ++i) base |= ((outcome >> (k - 1 - i)) & 1ull) << kb.bit[i];
    const uint64_t rest_mask = ((1ull << n) - 1) & ~kmask;
    const uint64_t NR = 1ull << (n - k);
    uint64_t g = (NR + 255) / 256;
    if (g > 65536) g = 65536;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SVH_C64)
        k_gather_rest<float><<<(unsigned)g, 256, 0, st>>>((const cpx<float> *)psi, base, rest_mask, NR,
                                                         (cpx<float> *)out);
    else
        k_gather_rest<double><<<(unsigned)g, 256, 0, st>>>((const cpx<double> *)psi, base, rest_mask, NR,
                                                          (cpx<double> *)out);
    return checked_launch("svh_collapse_gather");
}

int svh_partial_overlap(const void *a, const void *b, int n, int dtype, const int *kept, int k, void *out,
                        void *scratch, size_t scratch_bytes, void *stream) {
    if (!a || !b || !out) return fail(-1, "state or output pointer is NULL");
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "dtype must be SVH_C64 or SVH_C128");
    if (k > 6) return fail(-2, "svh_partial_overlap keeps at most 6 qubits");
    KeptBits kb{};
    uint64_t kmask;
    if (int e = check_kept(n, kept, k, kb, kmask)) return e;
    const int G = overlap_blocks(n, k);
    const int P = 1 << (2 * k);
    const size_t need = (size_t)G * P * sizeof(double2);
    if (!scratch || scratch_bytes < need)
        return fail(-5, "scratch too small for svh_partial_overlap (need " + std::to_string(need) + " bytes)");
    const uint64_t rest_mask = ((1ull << n) - 1) & ~kmask;
    const uint64_t chunk = (1ull << (n - k)) / (uint64_t)G;
    hipStream_t st = (hipStream_t)stream;
    double2 *part = (double2 *)scratch;
#define SVH_OVL(RT, QV)                                                                                  \
    k_overlap<RT, QV><<<(unsigned)G, 256, 0, st>>>((const cpx<RT> *)a, (const cpx<RT> *)b, kb, k, rest_mask, \
                                                   chunk, part)
    const int Q = P >= 256 ? P / 256 : 1;
    if (dtype == SVH_C64) {
        if (Q == 1) SVH_OVL(float, 1);
        else if (Q == 4) SVH_OVL(float, 4);
        else SVH_OVL(float, 16);
    } else {
        if (Q == 1) SVH_OVL(double, 1);
        else if (Q == 4) SVH_OVL(double, 4);
        else SVH_OVL(double, 16);
    }
#undef SVH_OVL
    k_overlap_final<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(part, P, G, (double2 *)out);
    return checked_launch("svh_partial_overlap");
}

}  // extern "C"
