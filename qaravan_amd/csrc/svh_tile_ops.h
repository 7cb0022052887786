// Device building blocks of the fused tile passes: in-register gate kernels
// (compile-time register indices), op dispatch, stage layout helpers, and the
// pieces the per-program specialised kernels are assembled from.  Included by
// svh_tile.hip and by generated kernel sources.
#pragma once
#include "svh_common.h"

namespace svh {

// LDS swizzle: XOR the low 5 bits with the next 5.  Spreads lane groups whose
// tile-local bits skip the register bits across the LDS banks.  It is linear
// over GF(2), so swz(a | b) = swz(a) ^ swz(b) for disjoint a, b: the host stores
// swz(register offset) per stage and the kernel XORs in swz(thread offset).
__device__ __forceinline__ int lds_swz(int li) { return li ^ ((li >> 5) & 31); }

// ------------------------------------------------ in-register gate kernels
template <typename R, int RB, int A>
__device__ __forceinline__ void ap_mat1(cpx<R> *v, cpx<R> m0, cpx<R> m1, cpx<R> m2, cpx<R> m3) {
#pragma unroll
    for (int p = 0; p < (1 << (RB - 1)); ++p) {
        const int i0 = ins0<A>(p), i1 = i0 | (1 << A);
        const cpx<R> x0 = v[i0], x1 = v[i1];
        v[i0] = cfma(m0, x0, cmul(m1, x1));
        v[i1] = cfma(m2, x0, cmul(m3, x1));
    }
}

// exp(-iθP) = c·I - i s·P
template <typename R, int RB, int A>
__device__ __forceinline__ void ap_rot1(cpx<R> *v, int pauli, R c, R s) {
#pragma unroll
    for (int p = 0; p < (1 << (RB - 1)); ++p) {
        const int i0 = ins0<A>(p), i1 = i0 | (1 << A);
        const cpx<R> x0 = v[i0], x1 = v[i1];
        if (pauli == 1) {  // X: y0 = c x0 - i s x1
            v[i0] = cpx<R>{c * x0.x + s * x1.y, c * x0.y - s * x1.x};
            v[i1] = cpx<R>{c * x1.x + s * x0.y, c * x1.y - s * x0.x};
        } else if (pauli == 2) {  // Y: y0 = c x0 - s x1, y1 = s x0 + c x1
            v[i0] = cpx<R>{c * x0.x - s * x1.x, c * x0.y - s * x1.y};
            v[i1] = cpx<R>{s * x0.x + c * x1.x, s * x0.y + c * x1.y};
        } else {  // Z: y0 = (c - i s) x0, y1 = (c + i s) x1
            v[i0] = cpx<R>{c * x0.x + s * x0.y, c * x0.y - s * x0.x};
            v[i1] = cpx<R>{c * x1.x - s * x1.y, c * x1.y + s * x1.x};
        }
    }
}

// Σ Im(conj(λ)·(Pψ)) over this thread's pairs.  The 2^RB products of one
// thread are summed in the state precision (fused multiply-adds; for complex64
// the rounding error stays below 2^(RB+1) FP32 ulps of Σ|λ||ψ| over the
// thread's amplitudes, the size of the amplitudes' own rounding); every sum
// across threads is FP64.
template <typename R, int RB, int A>
__device__ __forceinline__ double gr_rot1(const cpx<R> *v, const cpx<R> *w, int pauli) {
    R acc = R(0);
#pragma unroll
    for (int p = 0; p < (1 << (RB - 1)); ++p) {
        const int i0 = ins0<A>(p), i1 = i0 | (1 << A);
        const cpx<R> x0 = v[i0], x1 = v[i1], l0 = w[i0], l1 = w[i1];
        if (pauli == 1) {
            acc = im_conj_mul_acc(im_conj_mul_acc(acc, l0, x1), l1, x0);
        } else if (pauli == 2) {  // (Pψ)0 = -i x1, (Pψ)1 = i x0
            acc = im_conj_mul_acc(im_conj_mul_acc(acc, l0, cpx<R>{x1.y, -x1.x}), l1, cpx<R>{-x0.y, x0.x});
        } else {  // (Pψ)1 = -x1
            acc = im_conj_mul_acc(im_conj_mul_acc(acc, l0, x0), l1, cpx<R>{-x1.x, -x1.y});
        }
    }
    return (double)acc;
}

template <typename R, int HI, int LO>
struct Quad {
    static __device__ __forceinline__ int base(int q) { return ins0<HI>(ins0<LO>(q)); }
};

template <typename R, int RB, int HI, int LO>
__device__ __forceinline__ void ap_mat2(cpx<R> *v, const cpx<R> (&m)[16]) {
#pragma unroll
    for (int q = 0; q < (1 << (RB - 2)); ++q) {
        const int i00 = Quad<R, HI, LO>::base(q);
        const int id[4] = {i00, i00 | (1 << LO), i00 | (1 << HI), i00 | (1 << HI) | (1 << LO)};
        const cpx<R> x[4] = {v[id[0]], v[id[1]], v[id[2]], v[id[3]]};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cpx<R> acc = cmul(m[4 * r], x[0]);
            acc = cfma(m[4 * r + 1], x[1], acc);
            acc = cfma(m[4 * r + 2], x[2], acc);
            acc = cfma(m[4 * r + 3], x[3], acc);
            v[id[r]] = acc;
        }
    }
}

// CX: TGT_HI -> target HI, control LO; otherwise target LO, control HI.
// (Both orientations are compile-time so every register index is constant.)
template <typename R, int RB, int HI, int LO, bool TGT_HI>
__device__ __forceinline__ void ap_cx_t(cpx<R> *v) {
#pragma unroll
    for (int q = 0; q < (1 << (RB - 2)); ++q) {
        const int i00 = Quad<R, HI, LO>::base(q);
        const int i11 = i00 | (1 << HI) | (1 << LO);
        constexpr int flip = TGT_HI ? (1 << LO) : (1 << HI);
        const int ia = i00 | flip;
        const cpx<R> t = v[ia];
        v[ia] = v[i11];
        v[i11] = t;
    }
}

template <typename R, int RB, int HI, int LO>
__device__ __forceinline__ void ap_cz(cpx<R> *v) {
#pragma unroll
    for (int q = 0; q < (1 << (RB - 2)); ++q) {
        const int i11 = Quad<R, HI, LO>::base(q) | (1 << HI) | (1 << LO);
        v[i11] = cneg(v[i11]);
    }
}

// P⊗P acting on a quad: returns (P⊗P x) in y.
template <typename R>
__device__ __forceinline__ void pp_apply(int pauli, const cpx<R> (&x)[4], cpx<R> (&y)[4]) {
    if (pauli == 1) {  // XX: |b⟩ -> |~b⟩
        y[0] = x[3]; y[1] = x[2]; y[2] = x[1]; y[3] = x[0];
    } else if (pauli == 2) {  // YY: (YYx)_00 = -x_11, _11 = -x_00, _01 = x_10, _10 = x_01
        y[0] = cneg(x[3]); y[1] = x[2]; y[2] = x[1]; y[3] = cneg(x[0]);
    } else {  // ZZ: diag(1,-1,-1,1)
        y[0] = x[0]; y[1] = cneg(x[1]); y[2] = cneg(x[2]); y[3] = x[3];
    }
}

template <typename R, int RB, int HI, int LO>
__device__ __forceinline__ void ap_rot2(cpx<R> *v, int pauli, R c, R s) {
#pragma unroll
    for (int q = 0; q < (1 << (RB - 2)); ++q) {
        const int i00 = Quad<R, HI, LO>::base(q);
        const int id[4] = {i00, i00 | (1 << LO), i00 | (1 << HI), i00 | (1 << HI) | (1 << LO)};
        const cpx<R> x[4] = {v[id[0]], v[id[1]], v[id[2]], v[id[3]]};
        cpx<R> y[4];
        pp_apply<R>(pauli, x, y);
#pragma unroll
        for (int r = 0; r < 4; ++r)  // c x - i s y
            v[id[r]] = cpx<R>{c * x[r].x + s * y[r].y, c * x[r].y - s * y[r].x};
    }
}

// As gr_rot1, for the 2-qubit Pauli products (same precision).
template <typename R, int RB, int HI, int LO>
__device__ __forceinline__ double gr_rot2(const cpx<R> *v, const cpx<R> *w, int pauli) {
    R acc = R(0);
#pragma unroll
    for (int q = 0; q < (1 << (RB - 2)); ++q) {
        const int i00 = Quad<R, HI, LO>::base(q);
        const int id[4] = {i00, i00 | (1 << LO), i00 | (1 << HI), i00 | (1 << HI) | (1 << LO)};
        const cpx<R> x[4] = {v[id[0]], v[id[1]], v[id[2]], v[id[3]]};
        cpx<R> y[4];
        pp_apply<R>(pauli, x, y);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc = im_conj_mul_acc(acc, w[id[r]], y[r]);
    }
    return (double)acc;
}

// ------------------------------------------------ compile-time dispatch
// Register-bit indices arrive at run time (uniform across the workgroup); each
// switch arm instantiates the gate with constant indices so v[] stays in VGPRs.
#define SVH_SWITCH1(a, CALL)                                   \
    switch (a) {                                               \
        case 0: CALL(0); break;                                \
        case 1: if constexpr (RB > 1) { CALL(1); } break;      \
        case 2: if constexpr (RB > 2) { CALL(2); } break;      \
        case 3: if constexpr (RB > 3) { CALL(3); } break;      \
        case 4: if constexpr (RB > 4) { CALL(4); } break;      \
        default: break;                                        \
    }
#define SVH_C2(H, L, CALL) \
    case (H)*8 + (L): if constexpr (RB > (H)) { CALL(H, L); } break;
#define SVH_SWITCH2(hi, lo, CALL)                                                   \
    switch ((hi)*8 + (lo)) {                                                        \
        SVH_C2(1, 0, CALL)                                                          \
        SVH_C2(2, 0, CALL) SVH_C2(2, 1, CALL)                                       \
        SVH_C2(3, 0, CALL) SVH_C2(3, 1, CALL) SVH_C2(3, 2, CALL)                    \
        SVH_C2(4, 0, CALL) SVH_C2(4, 1, CALL) SVH_C2(4, 2, CALL) SVH_C2(4, 3, CALL) \
        default: break;                                                             \
    }

// One DPP step: v + (v as seen through the lane pattern CTRL), rows outside
// ROW_MASK adding 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v) {
    return v + __builtin_amdgcn_update_dpp(0.0, v, CTRL, ROW_MASK, 0xf, false);
}

// Sum over the lanes of one wave, returned to every lane; W = active lanes (64,
// or the whole workgroup when it is smaller than a wave: 16 or 32).  Row
// operations of the DPP network, no LDS: lane pairs (quad_perm 1,0,3,2), quads
// (quad_perm 2,3,0,1), 8 (row_half_mirror), 16 (row_mirror); then rows 1 and 3
// add the last lane of the row before (row_bcast15) and rows 2 and 3 add lane
// 31 (row_bcast31), which leaves the total in lane W-1.  Fixed order, so the
// result does not vary from run to run.
template <int W>
__device__ __forceinline__ double wave_sum(double v) {
    static_assert(W == 16 || W == 32 || W == 64, "wave_sum: 16, 32 or 64 lanes");
    v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
    v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
    v = dpp_add<0x140, 0xf>(v);  // row_mirror
    if constexpr (W >= 32) v = dpp_add<0x142, 0xa>(v);  // row_bcast15
    if constexpr (W == 64) v = dpp_add<0x143, 0xc>(v);  // row_bcast31
    // readlane moves 32-bit integers: take the two halves of the double as they are
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), W - 1),
                            __builtin_amdgcn_readlane(__double2loint(v), W - 1));
}

// ---- several gradient terms of one stage summed over the wave together
// Lane L's partner across lane bit MASK (L ^ MASK), through the DPP row
// network for bits 0, 1 and 3 and a lane permute for bits 4 and 5.
template <int MASK>
__device__ __forceinline__ double lane_xor(double x) {
    if constexpr (MASK == 1) return __builtin_amdgcn_update_dpp(0.0, x, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    else if constexpr (MASK == 2) return __builtin_amdgcn_update_dpp(0.0, x, 0x4e, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    else if constexpr (MASK == 8) return __builtin_amdgcn_update_dpp(0.0, x, 0x128, 0xf, 0xf, false);  // row_ror:8
    else return __shfl_xor(x, MASK, 64);
}

// One transposed step over lane bit B: M values per lane become M/2.  A lane
// whose bit B is 0 keeps the lower half of the values and gives its partner
// (lane ^ 2^B) the upper half; a lane whose bit is 1 does the opposite; each
// adds the partner's copy of the half it keeps.  Bit 2 has no single DPP
// partner pattern: row_ror:4 reaches L + 4 and row_ror:12 reaches L - 4 inside
// the row of 16, and the lane takes the one that is L ^ 4.
template <int B, int M>
__device__ __forceinline__ void xpose_step(double *v, int lane) {
    const bool up = (lane >> B) & 1;
#pragma unroll
    for (int j = 0; j < M / 2; ++j) {
        const double keep = up ? v[j + M / 2] : v[j];
        const double send = up ? v[j] : v[j + M / 2];
        double recv;
        if constexpr (B == 2) {
            const double fwd = __builtin_amdgcn_update_dpp(0.0, send, 0x124, 0xf, 0xf, false);  // row_ror:4
            const double bwd = __builtin_amdgcn_update_dpp(0.0, send, 0x12c, 0xf, 0xf, false);  // row_ror:12
            recv = up ? bwd : fwd;
        } else {
            recv = lane_xor<1 << B>(send);
        }
        v[j] = keep + recv;
    }
}

// Sums each of the 2^K values (K <= 3) of every lane over the W lanes of its
// wave (W = 16, 32 or 64: the workgroup when it is smaller than a wave).
// Transposed steps over lane bits 0..K-1 (2^K - 1 exchanges instead of 6·2^K),
// then plain sums over the other lane bits.  On return every lane holds the
// wave total of value rev_K(lane mod 2^K) (its low K lane bits reversed).
// Every sum is FP64 in a fixed order, so results do not vary from run to run.
template <int K, int W>
__device__ __forceinline__ double wave_sum_many(double (&v)[1 << K]) {
    static_assert(K >= 0 && K <= 3, "wave_sum_many: at most 8 values");
    static_assert(W == 16 || W == 32 || W == 64, "wave_sum_many: 16, 32 or 64 lanes");
    const int lane = threadIdx.x & 63;
    if constexpr (K >= 1) xpose_step<0, (1 << K)>(v, lane);
    if constexpr (K >= 2) xpose_step<1, (1 << K) / 2>(v, lane);
    if constexpr (K >= 3) xpose_step<2, 2>(v, lane);
    double s = v[0];
    if constexpr (K < 1) s += lane_xor<1>(s);
    if constexpr (K < 2) s += lane_xor<2>(s);
    if constexpr (K < 3) {
        // bits 2 and 3 together: lanes L, L + 4, L + 8, L + 12 of the row
        s += __builtin_amdgcn_update_dpp(0.0, s, 0x124, 0xf, 0xf, false);  // row_ror:4
        s += __builtin_amdgcn_update_dpp(0.0, s, 0x128, 0xf, 0xf, false);  // row_ror:8
    } else {
        s += lane_xor<8>(s);
    }
    if constexpr (W >= 32) s += lane_xor<16>(s);
    if constexpr (W == 64) s += lane_xor<32>(s);
    return s;
}

// Index of the value a lane holds after wave_sum_many<K>: its low K bits reversed.
template <int K>
__device__ __forceinline__ int rev_low_bits(int lane) {
    int j = 0;
#pragma unroll
    for (int t = 0; t < K; ++t) j |= ((lane >> t) & 1) << (K - 1 - t);
    return j;
}

// Store 2·(wave sum) of the first `count` of the 2^K gradient terms g into
// gacc[wave][slot0 + j] (LDS, one row of ng slots per wave).
template <int K, int W>
__device__ __forceinline__ void sp_grad_group(double (&g)[1 << K], double *gacc, int ng, int slot0, int count) {
    const double s = wave_sum_many<K, W>(g);
    const int lane = threadIdx.x & 63;
    const int j = rev_low_bits<K>(lane);
    if (lane < (1 << K) && j < count) gacc[(threadIdx.x >> 6) * ng + slot0 + j] = 2.0 * s;
}

template <typename R>
__device__ __forceinline__ cpx<R> ld_m(const svh_op_t &op, int i) {
    if constexpr (std::is_same<R, float>::value) return cpx<R>{op.mf[2 * i], op.mf[2 * i + 1]};
    else return cpx<R>{op.m[2 * i], op.m[2 * i + 1]};
}

// (cos θ, sin θ) of the op's angle from the per-call table (FP64-evaluated).
template <typename R>
__device__ __forceinline__ void op_sincos(const cpx<R> *trig, const svh_op_t &op, R &cs, R &sn) {
    const cpx<R> t = trig[op.param];
    cs = t.x;
    sn = (op.flags & SVH_OPF_NEG) ? -t.y : t.y;
}

// Apply one op to the registers (v: ψ, w: λ when DUAL).  Returns the gradient
// partial (this thread's Σ Im(conj λ · Pψ)) through *g when the op asks for it.
template <typename R, int RB, bool DUAL>
__device__ __forceinline__ void apply_op(cpx<R> (&v)[1 << RB], cpx<R> (&w)[DUAL ? (1 << RB) : 1],
                                         const svh_op_t &op, const cpx<R> *trig, double *g,
                                         bool *has_g) {
    const int kind = op.kind;
    if (kind == SVH_OP_MAT1) {
        const cpx<R> m0 = ld_m<R>(op, 0), m1 = ld_m<R>(op, 1), m2 = ld_m<R>(op, 2),
                     m3 = ld_m<R>(op, 3);
#define CALL_M1(A)                                      \
    ap_mat1<R, RB, A>(v, m0, m1, m2, m3);               \
    if constexpr (DUAL) ap_mat1<R, RB, A>(w, m0, m1, m2, m3);
        SVH_SWITCH1(op.a, CALL_M1)
#undef CALL_M1
    } else if (kind == SVH_OP_ROT1) {
        R cs, sn;
        op_sincos<R>(trig, op, cs, sn);
        const int pauli = op.pauli;
        if constexpr (DUAL) {
            if (op.flags & SVH_OPF_GRAD) {
                *has_g = true;
#define CALL_G1(A) *g = gr_rot1<R, RB, A>(v, w, pauli);
                SVH_SWITCH1(op.a, CALL_G1)
#undef CALL_G1
            }
        }
#define CALL_R1(A)                                   \
    ap_rot1<R, RB, A>(v, pauli, cs, sn);             \
    if constexpr (DUAL) ap_rot1<R, RB, A>(w, pauli, cs, sn);
        SVH_SWITCH1(op.a, CALL_R1)
#undef CALL_R1
    } else if (kind == SVH_OP_CX) {
        const int hi = op.a > op.b ? op.a : op.b, lo = op.a > op.b ? op.b : op.a;
        if (op.a > op.b) {
#define CALL_CXH(H, L)                               \
    ap_cx_t<R, RB, H, L, true>(v);                   \
    if constexpr (DUAL) ap_cx_t<R, RB, H, L, true>(w);
            SVH_SWITCH2(hi, lo, CALL_CXH)
#undef CALL_CXH
        } else {
#define CALL_CXL(H, L)                               \
    ap_cx_t<R, RB, H, L, false>(v);                  \
    if constexpr (DUAL) ap_cx_t<R, RB, H, L, false>(w);
            SVH_SWITCH2(hi, lo, CALL_CXL)
#undef CALL_CXL
        }
    } else if (kind == SVH_OP_CZ) {
        const int hi = op.a > op.b ? op.a : op.b, lo = op.a > op.b ? op.b : op.a;
#define CALL_CZ(H, L)                 \
    ap_cz<R, RB, H, L>(v);            \
    if constexpr (DUAL) ap_cz<R, RB, H, L>(w);
        SVH_SWITCH2(hi, lo, CALL_CZ)
#undef CALL_CZ
    } else if (kind == SVH_OP_MAT2) {
        cpx<R> m[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = ld_m<R>(op, i);
#define CALL_M2(H, L)                    \
    ap_mat2<R, RB, H, L>(v, m);          \
    if constexpr (DUAL) ap_mat2<R, RB, H, L>(w, m);
        SVH_SWITCH2(op.a, op.b, CALL_M2)
#undef CALL_M2
    } else if (kind == SVH_OP_ROT2) {
        R cs, sn;
        op_sincos<R>(trig, op, cs, sn);
        const int pauli = op.pauli;
        if constexpr (DUAL) {
            if (op.flags & SVH_OPF_GRAD) {
                *has_g = true;
#define CALL_G2(H, L) *g = gr_rot2<R, RB, H, L>(v, w, pauli);
                SVH_SWITCH2(op.a, op.b, CALL_G2)
#undef CALL_G2
            }
        }
#define CALL_R2(H, L)                             \
    ap_rot2<R, RB, H, L>(v, pauli, cs, sn);       \
    if constexpr (DUAL) ap_rot2<R, RB, H, L>(w, pauli, cs, sn);
        SVH_SWITCH2(op.a, op.b, CALL_R2)
#undef CALL_R2
    }
}

// Tile-local index of this thread's register 0 in a stage's layout: the thread
// index with zeros inserted at the (ascending) register bits.
template <int RB = 4>
__device__ __forceinline__ int thr_local(const svh_stage_t &S, int tid) {
    int t = tid;
#pragma unroll
    for (int k = 0; k < RB; ++k) {
        const int r = S.rsorted[k];
        t = ((t >> r) << (r + 1)) | (t & ((1 << r) - 1));
    }
    return t;
}

// Global amplitude offset of this thread's register 0 (thread bits only).
template <int TB>
__device__ __forceinline__ uint64_t thr_global(const svh_stage_t &S, int tid) {
    uint64_t a = 0;
#pragma unroll
    for (int i = 0; i < TB; ++i) a |= (uint64_t)((tid >> i) & 1) << S.tgbit[i];
    return a;
}

// ------------------------------------------------ specialised-pass pieces
// A program can be compiled into one kernel per pass in which every stage
// change and every op is spelled out with constant register indices (no
// run-time dispatch).  These helpers are the shared prologue / exchange /
// epilogue; matrices, angles and address tables still come from the program
// records in device memory, so one compiled program serves any θ.  RB is the
// number of register bits per thread (2^RB amplitudes in registers).
template <typename R, int TB, int RB = 4>
__device__ __forceinline__ uint64_t sp_base_wg(const svh_pass_t *__restrict__ pass, int n, uint64_t wg) {
    constexpr int L = TB + RB;
    uint64_t tmask = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) tmask |= 1ull << pass->bits[i];
    uint64_t base = 0, g = wg;
    for (int b = 0; b < n && g; ++b) {
        if (!((tmask >> b) & 1)) {
            base |= (g & 1ull) << b;
            g >>= 1;
        }
    }
    return base;
}

// Index of this workgroup's item in a batch table: the last item whose
// wg_begin is <= blockIdx.x (wg_begin ascending, first item at 0).
__device__ __forceinline__ int sp_find_item(const svh_batch_item_t *__restrict__ items, int nitems) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].wg_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename R, int TB, int RB = 4>
__device__ __forceinline__ void sp_load(const cpx<R> *__restrict__ psi, uint64_t base, const svh_stage_t &S,
                                        cpx<R> (&v)[1 << RB]) {
    const cpx<R> *p = psi + (base | thr_global<TB>(S, threadIdx.x));
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) v[j] = p[S.goff[j]];
}

// |0…0⟩ in the registers of the first layout, in place of sp_load: only
// amplitude 0 (base, thread part and register offset all zero; their bits are
// disjoint) is 1.
template <typename R, int TB, int RB = 4>
__device__ __forceinline__ void sp_gen_zero(uint64_t base, const svh_stage_t &S, cpx<R> (&v)[1 << RB]) {
    const uint64_t t = base | thr_global<TB>(S, threadIdx.x);
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) v[j] = cpx<R>{(t | (uint64_t)S.goff[j]) == 0 ? (R)1 : (R)0, (R)0};
}

template <typename R, int TB, int RB = 4>
__device__ __forceinline__ void sp_store(cpx<R> *__restrict__ psi, uint64_t base, const svh_stage_t &S,
                                         const cpx<R> (&v)[1 << RB]) {
    cpx<R> *p = psi + (base | thr_global<TB>(S, threadIdx.x));
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) p[S.goff[j]] = v[j];
}

// Re-distribute from layout P to layout S through LDS (FIRST: no previous
// exchange is still being read, so the leading barrier is skipped).
template <typename R, bool FIRST, int RB = 4>
__device__ __forceinline__ void sp_exchange(cpx<R> *lds, const svh_stage_t &P, const svh_stage_t &S,
                                            cpx<R> (&v)[1 << RB]) {
    const int tp = lds_swz(thr_local<RB>(P, threadIdx.x));
    if (!FIRST) __syncthreads();
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) lds[tp ^ P.lsw[j]] = v[j];
    __syncthreads();
    const int tn = lds_swz(thr_local<RB>(S, threadIdx.x));
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) v[j] = lds[tn ^ S.lsw[j]];
}

template <typename R, int A, int RB = 4>
__device__ __forceinline__ void sp_mat1(cpx<R> (&v)[1 << RB], const svh_op_t &op) {
    ap_mat1<R, RB, A>(v, ld_m<R>(op, 0), ld_m<R>(op, 1), ld_m<R>(op, 2), ld_m<R>(op, 3));
}
template <typename R, int H, int L, int RB = 4>
__device__ __forceinline__ void sp_mat2(cpx<R> (&v)[1 << RB], const svh_op_t &op) {
    cpx<R> m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = ld_m<R>(op, i);
    ap_mat2<R, RB, H, L>(v, m);
}
template <typename R, int A, int PAULI, int RB = 4>
__device__ __forceinline__ void sp_rot1(cpx<R> (&v)[1 << RB], const svh_op_t &op, const cpx<R> *trig) {
    R cs, sn;
    op_sincos<R>(trig, op, cs, sn);
    ap_rot1<R, RB, A>(v, PAULI, cs, sn);
}
template <typename R, int H, int L, int PAULI, int RB = 4>
__device__ __forceinline__ void sp_rot2(cpx<R> (&v)[1 << RB], const svh_op_t &op, const cpx<R> *trig) {
    R cs, sn;
    op_sincos<R>(trig, op, cs, sn);
    ap_rot2<R, RB, H, L>(v, PAULI, cs, sn);
}

// ---- adjoint (ψ and λ together) pieces
template <typename R, int TB, int RB = 4>
__device__ __forceinline__ void sp_load2(const cpx<R> *__restrict__ psi, const cpx<R> *__restrict__ lam,
                                         uint64_t base, const svh_stage_t &S, cpx<R> (&v)[1 << RB],
                                         cpx<R> (&w)[1 << RB]) {
    const uint64_t t = base | thr_global<TB>(S, threadIdx.x);
    const cpx<R> *p = psi + t;
    const cpx<R> *q = lam + t;
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) {
        v[j] = p[S.goff[j]];
        w[j] = q[S.goff[j]];
    }
}

template <typename R, int TB, int RB = 4>
__device__ __forceinline__ void sp_store2(cpx<R> *__restrict__ psi, cpx<R> *__restrict__ lam, uint64_t base,
                                          const svh_stage_t &S, const cpx<R> (&v)[1 << RB],
                                          const cpx<R> (&w)[1 << RB]) {
    const uint64_t t = base | thr_global<TB>(S, threadIdx.x);
    cpx<R> *p = psi + t;
    cpx<R> *q = lam + t;
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) {
        p[S.goff[j]] = v[j];
        q[S.goff[j]] = w[j];
    }
}

template <typename R, bool FIRST, int RB = 4>
__device__ __forceinline__ void sp_exchange2(cpx<R> *lds, cpx<R> *lds2, const svh_stage_t &P,
                                             const svh_stage_t &S, cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB]) {
    const int tp = lds_swz(thr_local<RB>(P, threadIdx.x));
    if (!FIRST) __syncthreads();
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) {
        lds[tp ^ P.lsw[j]] = v[j];
        lds2[tp ^ P.lsw[j]] = w[j];
    }
    __syncthreads();
    const int tn = lds_swz(thr_local<RB>(S, threadIdx.x));
#pragma unroll
    for (int j = 0; j < (1 << RB); ++j) {
        v[j] = lds[tn ^ S.lsw[j]];
        w[j] = lds2[tn ^ S.lsw[j]];
    }
}

// gacc[wave][slot] = 2·Σ_wave Im⟨λ|P|ψ⟩ (W = lanes per wave in this workgroup)
template <int W>
__device__ __forceinline__ void sp_grad_store(double g, double *gacc, int ng, int slot) {
    g = wave_sum<W>(g);
    if ((threadIdx.x & 63) == 0) gacc[(threadIdx.x >> 6) * ng + slot] = 2.0 * g;
}

template <typename R, int TB, int A, int PAULI, bool GRAD, int RB = 4>
__device__ __forceinline__ void spd_rot1(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t &op,
                                         const cpx<R> *trig, double *gacc, int ng) {
    if constexpr (GRAD) sp_grad_store<((1 << TB) < 64 ? (1 << TB) : 64)>(gr_rot1<R, RB, A>(v, w, PAULI), gacc, ng,
                                                                       op.grad_slot);
    R cs, sn;
    op_sincos<R>(trig, op, cs, sn);
    ap_rot1<R, RB, A>(w, PAULI, cs, sn);
    ap_rot1<R, RB, A>(v, PAULI, cs, sn);
}

template <typename R, int TB, int H, int L, int PAULI, bool GRAD, int RB = 4>
__device__ __forceinline__ void spd_rot2(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t &op,
                                         const cpx<R> *trig, double *gacc, int ng) {
    if constexpr (GRAD) sp_grad_store<((1 << TB) < 64 ? (1 << TB) : 64)>(gr_rot2<R, RB, H, L>(v, w, PAULI), gacc,
                                                                       ng, op.grad_slot);
    R cs, sn;
    op_sincos<R>(trig, op, cs, sn);
    ap_rot2<R, RB, H, L>(w, PAULI, cs, sn);
    ap_rot2<R, RB, H, L>(v, PAULI, cs, sn);
}

// Sum the waves' gradient slots of this workgroup into its partial row wg.
template <int TB>
__device__ __forceinline__ void sp_grad_flush_wg(const double *gacc, int ng, double *gpart, uint64_t wg) {
    constexpr int NW = ((1 << TB) + 63) / 64;
    __syncthreads();
    for (int sl = threadIdx.x; sl < ng; sl += (1 << TB)) {
        double acc = 0.0;
        for (int wv = 0; wv < NW; ++wv) acc += gacc[wv * ng + sl];
        gpart[wg * ng + sl] = acc;
    }
}

}  // namespace svh
