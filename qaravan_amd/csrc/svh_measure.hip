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

// Up to 16 outcomes: threads stream the whole state in address order (fully
// coalesced) and add |ψ|² into register bins selected by the kept bits; then
// one fixed-order tree per bin.  part[blockIdx.x*M + m].
template <typename R, int M>
__global__ void __launch_bounds__(256)
k_marg_reg(const cpx<R> *__restrict__ psi, uint64_t N, KeptBits kb, int k, double *__restrict__ part) {
    double acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0.0;
    // four independent loads in flight per thread per step
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t a = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    for (; a + 3 * stride < N; a += 4 * stride) {
        cpx<R> z[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) z[u] = psi[a + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t au = a + u * stride;
            const double p = abs2(z[u]);
            int bin = 0;
            for (int i = 0; i < k; ++i) bin = (bin << 1) | (int)((au >> kb.bit[i]) & 1);
#pragma unroll
            for (int m = 0; m < M; ++m) acc[m] += bin == m ? p : 0.0;
        }
    }
    for (; a < N; a += stride) {
        const double p = abs2(psi[a]);
        int bin = 0;
        for (int i = 0; i < k; ++i) bin = (bin << 1) | (int)((a >> kb.bit[i]) & 1);
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] += bin == m ? p : 0.0;
    }
    __shared__ double red[256];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        red[threadIdx.x] = acc[m];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(uint64_t)blockIdx.x * M + m] = red[0];
        __syncthreads();
    }
}

// Every qubit kept in address order (kept[i] = n-1-i): p[m] = |ψ[m]|².
template <typename R>
__global__ void __launch_bounds__(256) k_marg_full(const cpx<R> *__restrict__ psi, uint64_t N,
                                                   double *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t a = (uint64_t)blockIdx.x * 256 + threadIdx.x; a < N; a += stride) out[a] = abs2(psi[a]);
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
// Few kept qubits (M = 2^k ≤ 4): thread r walks rest indices in order (lanes on
// consecutive r, so each of the 2M loads per r is coalesced), keeping all M²
// products in FP64 registers; then one fixed-order tree per pair.
template <typename R, int M>
__global__ void __launch_bounds__(256)
k_overlap_reg(const cpx<R> *__restrict__ a, const cpx<R> *__restrict__ b, KeptBits kb, int k,
              uint64_t rest_mask, uint64_t NR, double2 *__restrict__ part) {
    constexpr int P = M * M;
    double2 acc[P];
#pragma unroll
    for (int p = 0; p < P; ++p) acc[p] = make_double2(0.0, 0.0);
    uint64_t off[M];
#pragma unroll
    for (int i = 0; i < M; ++i) off[i] = deposit_kept((uint64_t)i, kb, k);
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < NR) {
        uint64_t x = deposit_mask(r, rest_mask);
        const uint64_t d = deposit_mask(stride, rest_mask);
        for (; r < NR; r += stride) {
            cpx<R> u[M], w[M];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                u[i] = a[off[i] | x];
                w[i] = b[off[i] | x];
            }
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    acc[i * M + j].x += (double)u[i].x * w[j].x + (double)u[i].y * w[j].y;
                    acc[i * M + j].y += (double)u[i].y * w[j].x - (double)u[i].x * w[j].y;
                }
            x = masked_add(x, d, rest_mask);
        }
    }
    __shared__ double2 red[256];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        red[threadIdx.x] = acc[p];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                red[threadIdx.x].x += red[threadIdx.x + s].x;
                red[threadIdx.x].y += red[threadIdx.x + s].y;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(uint64_t)blockIdx.x * P + p] = red[0];
        __syncthreads();
    }
}

// More kept qubits (M = 2^k, 8 ≤ M ≤ 64): block g walks its rest chunk in
// batches of RBT = 1024/M indices; the batch's a(:, r) and b(:, r) are loaded
// with lanes on consecutive r (coalesced) into LDS as FP64 (rows padded by one
// element so the strided reads below hit distinct banks).  With M² ≥ 256 each
// thread owns a QI x QJ block of pairs (rows i0.., columns j0..), so every LDS
// value it reads feeds QI or QJ products; with M² < 256, T = 256/M² threads
// share a pair and split the batch.  part[g*P + i*M + j].
template <typename R, int M>
__global__ void __launch_bounds__(256)
k_overlap_lds(const cpx<R> *__restrict__ a, const cpx<R> *__restrict__ b, KeptBits kb, int k,
              uint64_t rest_mask, uint64_t chunk, double2 *__restrict__ part) {
    constexpr int P = M * M;
    constexpr int RBT = 1024 / M;              // rest indices per batch
    constexpr int LD = RBT + 1;                // padded LDS row
    constexpr int LPT = M * RBT / 256;         // loads per thread per vector (= 4)
    constexpr int Q = P >= 256 ? P / 256 : 1;
    constexpr int QI = Q >= 16 ? 4 : (Q >= 4 ? 2 : 1);
    constexpr int QJ = Q / QI;
    constexpr int T = P >= 256 ? 1 : 256 / P;
    __shared__ double2 sa[M * LD], sb[M * LD];
    const int t = threadIdx.x;
    const int rr = t % RBT;                    // this thread's rest slot in a batch
    uint64_t off[LPT];
#pragma unroll
    for (int l = 0; l < LPT; ++l) off[l] = deposit_kept((uint64_t)(t / RBT + l * (256 / RBT)), kb, k);
    const uint64_t drr = deposit_mask((uint64_t)rr, rest_mask);
    const uint64_t dbatch = deposit_mask((uint64_t)RBT, rest_mask);
    const int sub = P >= 256 ? 0 : t / P;
    // pair block: rows i0..i0+QI-1, columns j0..j0+QJ-1
    const int i0 = P >= 256 ? (t / (M / QJ)) * QI : (t % P) / M;
    const int j0 = P >= 256 ? (t % (M / QJ)) * QJ : (t % P) % M;
    double2 acc[QI][QJ];
#pragma unroll
    for (int x = 0; x < QI; ++x)
#pragma unroll
        for (int y = 0; y < QJ; ++y) acc[x][y] = make_double2(0.0, 0.0);
    const uint64_t r0 = (uint64_t)blockIdx.x * chunk;
    uint64_t xb = deposit_mask(r0, rest_mask);
    const int nb = (int)((chunk + RBT - 1) / RBT);
    for (int bt = 0; bt < nb; ++bt) {
        const bool live = (uint64_t)bt * RBT + rr < chunk;
        const uint64_t x = masked_add(xb, drr, rest_mask);
#pragma unroll
        for (int l = 0; l < LPT; ++l) {
            const int i = t / RBT + l * (256 / RBT);
            cpx<R> u = live ? a[off[l] | x] : cpx<R>{(R)0, (R)0};
            cpx<R> w = live ? b[off[l] | x] : cpx<R>{(R)0, (R)0};
            sa[i * LD + rr] = make_double2((double)u.x, (double)u.y);
            sb[i * LD + rr] = make_double2((double)w.x, (double)w.y);
        }
        __syncthreads();
        for (int c = sub; c < RBT; c += T) {
            double2 u[QI], w[QJ];
#pragma unroll
            for (int x = 0; x < QI; ++x) u[x] = sa[(i0 + x) * LD + c];
#pragma unroll
            for (int y = 0; y < QJ; ++y) w[y] = sb[(j0 + y) * LD + c];
#pragma unroll
            for (int x = 0; x < QI; ++x)
#pragma unroll
                for (int y = 0; y < QJ; ++y) {
                    acc[x][y].x += u[x].x * w[y].x + u[x].y * w[y].y;
                    acc[x][y].y += u[x].y * w[y].x - u[x].x * w[y].y;
                }
        }
        __syncthreads();
        xb = masked_add(xb, dbatch, rest_mask);
    }
    if (P >= 256) {
#pragma unroll
        for (int x = 0; x < QI; ++x)
#pragma unroll
            for (int y = 0; y < QJ; ++y) part[(uint64_t)blockIdx.x * P + (i0 + x) * M + (j0 + y)] = acc[x][y];
        return;
    }
    __shared__ double2 red[256];
    red[t] = acc[0][0];
    __syncthreads();
    for (int s2 = T / 2; s2 > 0; s2 >>= 1) {
        if (sub < s2) {
            const double2 o = red[t + s2 * P];
            red[t].x += o.x;
            red[t].y += o.y;
        }
        __syncthreads();
    }
    if (sub == 0) part[(uint64_t)blockIdx.x * P + t] = red[t];
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
    int kind;            // 0: register bins (k <= 4), 1: block per (m, chunk), 2: thread per (m, chunk)
    uint64_t M, C, chunk;
};

static MargPlan marg_plan(int n, int k) {
    MargPlan p;
    p.M = 1ull << k;
    const uint64_t N = 1ull << n;
    const uint64_t NR = 1ull << (n - k);
    if (k <= 4) {
        p.kind = 0;
        uint64_t g = (N + 255) / 256;
        p.C = g < 1024 ? g : 1024;
        p.chunk = 0;
        return p;
    }
    if (p.M <= 4096) {
        p.kind = 1;
        uint64_t C = 4096 / p.M;
        const uint64_t cap = NR >= 1024 ? NR / 1024 : 1;
        if (C > cap) C = cap;
        p.C = C;
    } else {
        p.kind = 2;
        uint64_t C = p.M >= (1ull << 18) ? 1 : (1ull << 18) / p.M;
        if (C > NR) C = NR;
        p.C = C;
    }
    p.chunk = NR / p.C;
    return p;
}

static size_t scan_blocks(uint64_t M) { return (size_t)((M + kScanTile - 1) / kScanTile); }

struct OverlapPlan {
    bool reg;            // k <= 2: k_overlap_reg (grid-stride); else k_overlap_lds (chunks)
    int G;
    uint64_t chunk;
};

static OverlapPlan overlap_plan(int n, int k) {
    OverlapPlan o;
    const uint64_t P = 1ull << (2 * k);
    const uint64_t NR = 1ull << (n - k);
    if (k <= 2) {
        o.reg = true;
        const uint64_t g = (NR + 255) / 256;
        o.G = (int)(g < 1024 ? g : 1024);
        o.chunk = 0;
        return o;
    }
    o.reg = false;
    const uint64_t rbt = 1024ull >> k;
    uint64_t G = (1ull << 18) / P;
    if (G > 1024) G = 1024;
    const uint64_t cap = NR >= rbt ? NR / rbt : 1;
    if (G > cap) G = cap;
    if (G < 1) G = 1;
    o.G = (int)G;
    o.chunk = NR / G;
    return o;
}

}  // namespace svh

using namespace svh;

extern "C" {

size_t svh_measure_scratch_bytes(int what, int n, int k) {
    if (n < 1 || n > kMaxQubits || k < 0 || k > n) return 0;
    switch (what) {
    case SVH_MEAS_PROBS: {
        const MargPlan p = marg_plan(n, k);
        return p.kind == 0 || p.C > 1 ? (size_t)(p.C * p.M * sizeof(double)) : sizeof(double);
    }
    case SVH_MEAS_SAMPLE: {
        const uint64_t M = 1ull << k;
        return (size_t)(M * sizeof(double) + 2 * scan_blocks(M) * sizeof(double));
    }
    case SVH_MEAS_OVERLAP:
        if (k > 6) return 0;
        return (size_t)overlap_plan(n, k).G * ((size_t)1 << (2 * k)) * sizeof(double2);
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
    const bool staged = p.kind == 0 || p.C > 1;
    if (staged && (!scratch || scratch_bytes < need))
        return fail(-5, "scratch too small for svh_marginal_probs (need " + std::to_string(need) + " bytes)");
    const uint64_t rest_mask = ((1ull << n) - 1) & ~kmask;
    hipStream_t st = (hipStream_t)stream;
    double *part = staged ? (double *)scratch : probs;
    bool natural = k == n;
    for (int i = 0; i < k && natural; ++i) natural = kb.bit[i] == n - 1 - i;
    if (natural && k > 4) {
        const uint64_t N = 1ull << n;
        uint64_t g = (N + 255) / 256;
        if (g > 65536) g = 65536;
        if (dtype == SVH_C64)
            k_marg_full<float><<<(unsigned)g, 256, 0, st>>>((const cpx<float> *)psi, N, probs);
        else
            k_marg_full<double><<<(unsigned)g, 256, 0, st>>>((const cpx<double> *)psi, N, probs);
        return checked_launch("svh_marginal_probs");
    }
    if (p.kind == 0) {
        const unsigned g = (unsigned)p.C;
        const uint64_t N = 1ull << n;
#define SVH_MREG(RT, MV) \
    k_marg_reg<RT, MV><<<g, 256, 0, st>>>((const cpx<RT> *)psi, N, kb, k, part)
        if (dtype == SVH_C64) {
            if (p.M == 1) SVH_MREG(float, 1);
            else if (p.M == 2) SVH_MREG(float, 2);
            else if (p.M == 4) SVH_MREG(float, 4);
            else if (p.M == 8) SVH_MREG(float, 8);
            else SVH_MREG(float, 16);
        } else {
            if (p.M == 1) SVH_MREG(double, 1);
            else if (p.M == 2) SVH_MREG(double, 2);
            else if (p.M == 4) SVH_MREG(double, 4);
            else if (p.M == 8) SVH_MREG(double, 8);
            else SVH_MREG(double, 16);
        }
#undef SVH_MREG
    } else if (p.kind == 1) {
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
    if (staged) k_marg_final<<<(unsigned)((p.M + 255) / 256), 256, 0, st>>>(part, p.M, p.C, probs);
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
    const OverlapPlan o = overlap_plan(n, k);
    const int P = 1 << (2 * k);
    const size_t need = (size_t)o.G * P * sizeof(double2);
    if (!scratch || scratch_bytes < need)
        return fail(-5, "scratch too small for svh_partial_overlap (need " + std::to_string(need) + " bytes)");
    const uint64_t rest_mask = ((1ull << n) - 1) & ~kmask;
    const uint64_t NR = 1ull << (n - k);
    hipStream_t st = (hipStream_t)stream;
    double2 *part = (double2 *)scratch;
    const unsigned G = (unsigned)o.G;
#define SVH_OREG(RT, MV)                                                                                 \
    k_overlap_reg<RT, MV><<<G, 256, 0, st>>>((const cpx<RT> *)a, (const cpx<RT> *)b, kb, k, rest_mask, NR, \
                                             part)
#define SVH_OLDS(RT, MV)                                                                                 \
    k_overlap_lds<RT, MV><<<G, 256, 0, st>>>((const cpx<RT> *)a, (const cpx<RT> *)b, kb, k, rest_mask,    \
                                             o.chunk, part)
    if (dtype == SVH_C64) {
        switch (k) {
            case 0: SVH_OREG(float, 1); break;
            case 1: SVH_OREG(float, 2); break;
            case 2: SVH_OREG(float, 4); break;
            case 3: SVH_OLDS(float, 8); break;
            case 4: SVH_OLDS(float, 16); break;
            case 5: SVH_OLDS(float, 32); break;
            default: SVH_OLDS(float, 64); break;
        }
    } else {
        switch (k) {
            case 0: SVH_OREG(double, 1); break;
            case 1: SVH_OREG(double, 2); break;
            case 2: SVH_OREG(double, 4); break;
            case 3: SVH_OLDS(double, 8); break;
            case 4: SVH_OLDS(double, 16); break;
            case 5: SVH_OLDS(double, 32); break;
            default: SVH_OLDS(double, 64); break;
        }
    }
#undef SVH_OREG
#undef SVH_OLDS
    k_overlap_final<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(part, P, o.G, (double2 *)out);
    return checked_launch("svh_partial_overlap");
}

}  // extern "C"
