// Dense 5-qubit blocks on the FP32 matrix cores (complex64 states).
//
// A 32x32 complex unitary U = a + i·b acting on 5 address bits multiplies a
// 32 x C complex matrix X = x + i·y whose rows are the 32 block amplitudes and
// whose C columns run over the other bits.  It is done as three real FP32
// GEMMs (the 3M form): P1 = a·x, P2 = b·y, P3 = (a + b)·(x + y), then
// Re = P1 − P2 and Im = P3 − P1 − P2 — 3/4 of the MFMA work of the 64x64 real
// form [[a, −b], [b, a]].  One workgroup (4 waves) owns a tile of 4096
// amplitudes: the 5 block bits and the 7 lowest other bits, so every tile is a
// set of >= 128-amplitude contiguous runs in HBM.  The tile is staged in LDS
// with coalesced loads, each wave multiplies a 32-column slab with
// v_mfma_f32_32x32x2_f32 (a, b and a + b live in 48 VGPRs per lane for the
// whole launch), writes its result back to LDS, and the tile is stored
// coalesced.  The grid is persistent (workgroups stride over tiles) so the
// fragments are built once per WG.
//
// Operand layouts of v_mfma_f32_32x32x2_f32 (wave64):
//   A[i][k]: lane i + 32k;  B[k][j]: lane j + 32k;
//   D[i][j]: lane j + 32·((i/4)%2), register (i%4) + 4·(i/8).
#include "svh_common.h"

namespace svh {

typedef float floatx16 __attribute__((ext_vector_type(16)));

struct Block5Args {
    int64_t gk[16];   // global offset of tile element 256·k (k < 16)
    int32_t lk[16];   // LDS complex index of tile element 256·k
    int32_t tg[8];    // global bit of tile-order bit i (i < 8, the thread bits)
    int32_t tl[8];    // LDS index bit of tile-order bit i
    int64_t tmask;    // global bits inside the tile
    int64_t ntiles;   // 2^(n-12)
    int32_t n;
    int32_t pad;
};

// LDS layout of a tile: complex index r·128 + c (r = block row, c = column).
__global__ void __launch_bounds__(256)
k_block5(cpx<float> *__restrict__ psi, const cpx<float> *__restrict__ U, Block5Args A) {
    __shared__ cpx<float> tile[4096];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int j32 = lane & 31;
    const int hi = lane >> 5;  // 0: real rows / k even, 1: imaginary rows / k odd

    // 3M complex product (U = a + i·b, X = x + i·y):
    //   Re(UX) = a·x − b·y,  Im(UX) = (a + b)·(x + y) − a·x − b·y.
    // MFMA step t (k = 2t + hi): lane (j32, hi) holds U[j32][2t + hi] as the A
    // operand and amplitude 2t + hi of its column as the B operand.
    float Fr[16], Fi[16], Fs[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const cpx<float> u = U[j32 * 32 + 2 * t + hi];
        Fr[t] = u.x;
        Fi[t] = u.y;
        Fs[t] = u.x + u.y;
    }

    // this thread's share of the tile-order index: t = tid (8 bits)
    uint64_t gt = 0;
    int lt = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        gt |= (uint64_t)((tid >> i) & 1) << A.tg[i];
        lt |= ((tid >> i) & 1) << A.tl[i];
    }
    const int col = 32 * wave + j32;

    for (int64_t t = blockIdx.x; t < A.ntiles; t += gridDim.x) {
        // tile base: t deposited into the non-tile bits
        uint64_t base = 0;
        {
            uint64_t g = (uint64_t)t;
            for (int b = 0; b < A.n && g; ++b) {
                if (!((A.tmask >> b) & 1)) {
                    base |= (g & 1ull) << b;
                    g >>= 1;
                }
            }
        }
        const cpx<float> *src = psi + (base | gt);
        cpx<float> vals[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) vals[k] = src[A.gk[k]];
        __syncthreads();  // previous tile's stores from LDS are done
#pragma unroll
        for (int k = 0; k < 16; ++k) tile[lt | A.lk[k]] = vals[k];
        __syncthreads();

        floatx16 p1 = {0}, p2 = {0}, p3 = {0};
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const cpx<float> x = tile[(2 * t + hi) * 128 + col];
            p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(Fr[t], x.x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f32_32x32x2f32(Fi[t], x.y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_32x32x2f32(Fs[t], x.x + x.y, p3, 0, 0, 0);
        }
        // a wave reads and writes only its own 32 columns, so no barrier
        // separates its reads and writes
        // D layout: row i = (v & 3) + 4 hi + 8 (v >> 2) of register v
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = (v & 3) + 4 * hi + 8 * (v >> 2);
            tile[i * 128 + col] = cpx<float>{p1[v] - p2[v], p3[v] - p1[v] - p2[v]};
        }
        __syncthreads();
        cpx<float> *dst = psi + (base | gt);
#pragma unroll
        for (int k = 0; k < 16; ++k) dst[A.gk[k]] = tile[lt | A.lk[k]];
    }
}

// ------------------------------------------------ two blocks per sweep
// Two 5-qubit unitaries A and B on disjoint bits (they commute) applied to
// one 4096-amplitude tile in one HBM sweep: twice the MFMA work per byte.
// Tile = A's 5 bits, B's 5 bits and the 2 lowest other bits; LDS complex index
// (before the swizzle) = rA·128 + rB·4 + o.  GEMM A runs over the 128 columns
// (rB, o), GEMM B over the 128 columns (rA, o).  The swizzle XORs rB's low
// three bits with rA's so the strided GEMM-B reads hit distinct banks while
// GEMM-A reads stay contiguous.
//
// When s of the address bits 0..3 lie outside the tile, 2^s tiles share every
// 128-byte line.  The persistent grid then makes those sibling tiles run at
// the same step on workgroups of the same XCD (workgroups are placed on the 8
// XCDs round-robin), so each line is fetched into that XCD's L2 once.
struct Block5x2Args {
    int64_t gk[16];   // global offset of tile element 256·k (k < 16)
    int32_t lk[16];   // LDS complex index of tile element 256·k
    int32_t tg[8];    // global bit of tile-order bit i (the thread bits)
    int32_t tl[8];    // LDS index bit of tile-order bit i
    int64_t tmask;    // global bits inside the tile
    int64_t ntiles;   // 2^(n-12)
    int32_t n;
    int32_t sib_bits; // s: sibling tiles grouped per XCD (see above), 0: plain order
};

__device__ __forceinline__ int swz2(int idx) { return idx ^ (((idx >> 7) & 7) << 2); }

__global__ void __launch_bounds__(256, 2)
k_block5x2(cpx<float> *__restrict__ psi, const cpx<float> *__restrict__ UA, const cpx<float> *__restrict__ UB,
           Block5x2Args A) {
    __shared__ cpx<float> tile[4096];
    // 3M complex products: with U = a + i·b and X = x + i·y,
    //   Re(UX) = a·x − b·y,  Im(UX) = (a + b)·(x + y) − a·x − b·y,
    // three real 32x32 GEMMs instead of the four of the real 64x64 form.
    // MFMA step t (k = 2t + hi): lane (j32, hi) holds U[j32][2t + hi] as the A
    // operand and amplitude 2t + hi of its column as the B operand.
    // A's fragments live in registers, B's in LDS: wb[t][lane] = (b_re, b_im, b_re + b_im).
    __shared__ float4 wb[16 * 64];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int j32 = lane & 31;
    const int hi = lane >> 5;

    float Fr[16], Fi[16], Fs[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const cpx<float> u = UA[j32 * 32 + 2 * t + hi];
        Fr[t] = u.x;
        Fi[t] = u.y;
        Fs[t] = u.x + u.y;
    }
    for (int e = tid; e < 16 * 64; e += 256) {
        const int t = e >> 6, l = e & 63;
        const cpx<float> w = UB[(l & 31) * 32 + 2 * t + (l >> 5)];
        wb[e] = make_float4(w.x, w.y, w.x + w.y, 0.f);
    }
    uint64_t gt = 0;
    int lt = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        gt |= (uint64_t)((tid >> i) & 1) << A.tg[i];
        lt |= ((tid >> i) & 1) << A.tl[i];
    }
    const int col = 32 * wave + j32;

    // tile order: tile = g·S + m with S = 2^s siblings; workgroup w takes
    // sibling m = (w / 8) mod S of the groups g = gw, gw + G/S, ... (gw = w
    // with those bits removed), so the S workgroups of one group are w, w + 8,
    // ..., on one XCD.  The grid is a multiple of 8·S.
    const int64_t G = gridDim.x;
    const int64_t w = blockIdx.x;
    const int sb = A.sib_bits;
    const int64_t S = 1ll << sb;
    const int64_t m = (w >> 3) & (S - 1);
    const int64_t gw = ((w >> (3 + sb)) << 3) | (w & 7);
    const int64_t gstride = G >> sb;
    for (int64_t g = gw; g < (A.ntiles >> sb); g += gstride) {
        // tile base: tile index deposited into the non-tile bits
        uint64_t base = 0;
        {
            uint64_t q = (uint64_t)(g * S + m);
            for (int b = 0; b < A.n && q; ++b) {
                if (!((A.tmask >> b) & 1)) {
                    base |= (q & 1ull) << b;
                    q >>= 1;
                }
            }
        }
        const cpx<float> *src = psi + (base | gt);
        cpx<float> vals[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) vals[k] = src[A.gk[k]];
        __syncthreads();  // the previous tile's stores have read LDS
#pragma unroll
        for (int k = 0; k < 16; ++k) tile[swz2(lt | A.lk[k])] = vals[k];
        __syncthreads();

        // GEMM A: rows rA, column col = (rB, o)
        floatx16 p1 = {0}, p2 = {0}, p3 = {0};
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const cpx<float> x = tile[swz2((2 * t + hi) * 128 + col)];
            p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(Fr[t], x.x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f32_32x32x2f32(Fi[t], x.y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_32x32x2f32(Fs[t], x.x + x.y, p3, 0, 0, 0);
        }
        // a wave reads and writes only its own 32 columns (the swizzle keeps
        // them in its own slots), so no barrier separates its reads and writes
        // D layout: row i = (v & 3) + 4 hi + 8 (v >> 2) of register v
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = (v & 3) + 4 * hi + 8 * (v >> 2);
            tile[swz2(i * 128 + col)] = cpx<float>{p1[v] - p2[v], p3[v] - p1[v] - p2[v]};
        }
        __syncthreads();

        // GEMM B: rows rB, column col = (rA, o): complex index rA·128 + rB·4 + o
        const int cb = (col >> 2) * 128 + (col & 3);
        floatx16 q1 = {0}, q2 = {0}, q3 = {0};
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const cpx<float> x = tile[swz2(cb + (2 * t + hi) * 4)];
            const float4 w = wb[t * 64 + lane];
            q1 = __builtin_amdgcn_mfma_f32_32x32x2f32(w.x, x.x, q1, 0, 0, 0);
            q2 = __builtin_amdgcn_mfma_f32_32x32x2f32(w.y, x.y, q2, 0, 0, 0);
            q3 = __builtin_amdgcn_mfma_f32_32x32x2f32(w.z, x.x + x.y, q3, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = (v & 3) + 4 * hi + 8 * (v >> 2);
            tile[swz2(cb + i * 4)] = cpx<float>{q1[v] - q2[v], q3[v] - q1[v] - q2[v]};
        }
        __syncthreads();
        cpx<float> *dst = psi + (base | gt);
#pragma unroll
        for (int k = 0; k < 16; ++k) dst[A.gk[k]] = tile[swz2(lt | A.lk[k])];
    }
}

}  // namespace svh

using namespace svh;

// Tile geometry for a block on address bits bits[0..5) (bits[j] = row bit j).
static int block5_args(int n, const int *bits, Block5Args *A) {
    if (n < 12) return fail(-2, "MFMA block path needs n >= 12");
    uint64_t bm = 0;
    for (int j = 0; j < 5; ++j) {
        if (bits[j] < 0 || bits[j] >= n || ((bm >> bits[j]) & 1)) return fail(-4, "bad block bits");
        bm |= 1ull << bits[j];
    }
    // tile = the 5 block bits + the 7 lowest other bits; LDS index bit of each
    int other[7], no = 0;
    for (int b = 0; b < n && no < 7; ++b)
        if (!((bm >> b) & 1)) other[no++] = b;
    uint64_t tmask = bm;
    for (int i = 0; i < 7; ++i) tmask |= 1ull << other[i];
    int tbits[12], nt = 0;
    for (int b = 0; b < n; ++b)
        if ((tmask >> b) & 1) tbits[nt++] = b;
    // LDS index of tile-order bit p: block row bit j -> 7 + j, other bit i -> i
    int lpos[12];
    for (int p = 0; p < 12; ++p) {
        const int g = tbits[p];
        lpos[p] = -1;
        for (int j = 0; j < 5; ++j)
            if (bits[j] == g) lpos[p] = 7 + j;
        for (int i = 0; i < 7; ++i)
            if (other[i] == g) lpos[p] = i;
    }
    for (int i = 0; i < 8; ++i) {
        A->tg[i] = tbits[i];
        A->tl[i] = lpos[i];
    }
    for (int k = 0; k < 16; ++k) {
        int64_t go = 0;
        int lo = 0;
        for (int i = 0; i < 4; ++i) {
            if ((k >> i) & 1) {
                go |= 1ll << tbits[8 + i];
                lo |= 1 << lpos[8 + i];
            }
        }
        A->gk[k] = go;
        A->lk[k] = lo;
    }
    A->tmask = (int64_t)tmask;
    A->ntiles = 1ll << (n - 12);
    A->n = n;
    A->pad = 0;
    return 0;
}

extern "C" int svh_apply_block5(void *psi, int n, int dtype, const int *bits, const void *u_dev,
                                void *stream) {
    if (!psi || !bits || !u_dev) return fail(-1, "NULL pointer");
    if (dtype != SVH_C64) return fail(-3, "the MFMA block path is complex64 only");
    if (n < 12 || n > kMaxQubits) return fail(-2, "MFMA block path needs 12 <= n <= 36");
    Block5Args A;
    if (int e = block5_args(n, bits, &A)) return e;
    int dev = 0, cus = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    if (int e = hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                           "hipDeviceGetAttribute"))
        return e;
    // two resident workgroups per CU (150 VGPRs + 32 AGPRs per lane)
    const int64_t cap = 2 * (int64_t)(cus > 0 ? cus : 256);
    const int64_t grid = A.ntiles < cap ? A.ntiles : cap;
    k_block5<<<(unsigned)grid, 256, 0, (hipStream_t)stream>>>((cpx<float> *)psi, (const cpx<float> *)u_dev, A);
    return checked_launch("k_block5");
}


// Tile geometry for blocks on address bits ba[0..5) and bb[0..5) (row bit j
// of each matrix): tile = both blocks + the 2 lowest other bits.
static int block5x2_args(int n, const int *ba, const int *bb, Block5x2Args *A) {
    if (n < 12) return fail(-2, "MFMA block path needs n >= 12");
    uint64_t ma = 0, mb = 0;
    for (int j = 0; j < 5; ++j) {
        if (ba[j] < 0 || ba[j] >= n || ((ma >> ba[j]) & 1)) return fail(-4, "bad block bits");
        if (bb[j] < 0 || bb[j] >= n || ((mb >> bb[j]) & 1)) return fail(-4, "bad block bits");
        ma |= 1ull << ba[j];
        mb |= 1ull << bb[j];
    }
    if (ma & mb) return fail(-4, "the two blocks share a bit");
    int other[2], no = 0;
    for (int b = 0; b < n && no < 2; ++b)
        if (!(((ma | mb) >> b) & 1)) other[no++] = b;
    const uint64_t tmask = ma | mb | (1ull << other[0]) | (1ull << other[1]);
    int tbits[12], nt = 0;
    for (int b = 0; b < n; ++b)
        if ((tmask >> b) & 1) tbits[nt++] = b;
    // LDS bit of tile-order bit p: A row bit j -> 7 + j, B row bit j -> 2 + j, other i -> i
    int lpos[12];
    for (int p = 0; p < 12; ++p) {
        const int g = tbits[p];
        lpos[p] = -1;
        for (int j = 0; j < 5; ++j) {
            if (ba[j] == g) lpos[p] = 7 + j;
            if (bb[j] == g) lpos[p] = 2 + j;
        }
        for (int i = 0; i < 2; ++i)
            if (other[i] == g) lpos[p] = i;
    }
    for (int i = 0; i < 8; ++i) {
        A->tg[i] = tbits[i];
        A->tl[i] = lpos[i];
    }
    for (int k = 0; k < 16; ++k) {
        int64_t go = 0;
        int lo = 0;
        for (int i = 0; i < 4; ++i) {
            if ((k >> i) & 1) {
                go |= 1ll << tbits[8 + i];
                lo |= 1 << lpos[8 + i];
            }
        }
        A->gk[k] = go;
        A->lk[k] = lo;
    }
    A->tmask = (int64_t)tmask;
    A->ntiles = 1ll << (n - 12);
    A->n = n;
    // siblings: the non-tile bits among 0..3 (at most two, as the tile holds
    // the two lowest non-block bits) are the lowest non-tile bits, so they are
    // the low bits of the tile index
    int sb = 0;
    for (int b = 0; b < 4; ++b) sb += !((tmask >> b) & 1);
    A->sib_bits = A->ntiles >= (8ll << sb) ? sb : 0;
    return 0;
}

extern "C" int svh_apply_block5x2(void *psi, int n, int dtype, const int *bits_a, const void *ua_dev,
                                  const int *bits_b, const void *ub_dev, void *stream) {
    if (!psi || !bits_a || !bits_b || !ua_dev || !ub_dev) return fail(-1, "NULL pointer");
    if (dtype != SVH_C64) return fail(-3, "the MFMA block path is complex64 only");
    if (n < 12 || n > kMaxQubits) return fail(-2, "MFMA block path needs 12 <= n <= 36");
    Block5x2Args A;
    if (int e = block5x2_args(n, bits_a, bits_b, &A)) return e;
    int dev = 0, cus = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    if (int e = hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                           "hipDeviceGetAttribute"))
        return e;
    // two resident workgroups per CU; with 2^s siblings the grid is a
    // multiple of 8·2^s (8 XCDs) and at most the tile count
    const int64_t quant = A.sib_bits ? 8ll << A.sib_bits : 1;
    int64_t grid = 2 * (int64_t)(cus > 0 ? cus : 256);
    grid = grid / quant * quant;
    if (grid > A.ntiles) grid = A.ntiles;  // ntiles >= quant and both are powers of two
    if (grid < 1) grid = 1;
    k_block5x2<<<(unsigned)grid, 256, 0, (hipStream_t)stream>>>((cpx<float> *)psi, (const cpx<float> *)ua_dev,
                                                                (const cpx<float> *)ub_dev, A);
    return checked_launch("k_block5x2");
}
