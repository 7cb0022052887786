// Generic tile pass: one kernel body for every pass structure.  The stage
// records are walked at run time; a stage whose op sequence is in the compiled
// pattern library (svh_patterns.h) runs as straight-line code with constant
// register indices, any other stage is interpreted op by op.  Used by the
// single-program kernels k_gtile and by the persistent batch kernel.
//
// Launch shape: 2^tb active threads per tile (tb >= 6, whole waves), 2^rb
// amplitudes per thread; a workgroup may be larger than 2^tb (the persistent
// batch kernel has one size for all its circuits): the extra waves take part
// in the barriers only.  In an adjoint sweep every active wave writes its own
// row of gradient partials: row (tile << (tb - 6)) + wave of n_grad doubles.
#pragma once
#include "svh_tile_ops.h"

namespace svh {

// Adjoint rotation: the wave's Σ 2·Im⟨λ|P|ψ⟩ goes to the wave's partial row
// (gout[slot], written by lane 0), then U† is applied to λ and ψ.
template <typename R, int RB, int A, int PAULI, bool GRAD>
__device__ __forceinline__ void gd_rot1(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t &op,
                                        const cpx<R> *trig, double *gout, bool lane0) {
    if constexpr (GRAD) {
        const double g = wave_sum<64>(gr_rot1<R, RB, A>(v, w, PAULI));
        if (lane0) gout[op.grad_slot] = 2.0 * g;
    }
    R cs, sn;
    op_sincos<R>(trig, op, cs, sn);
    ap_rot1<R, RB, A>(w, PAULI, cs, sn);
    ap_rot1<R, RB, A>(v, PAULI, cs, sn);
}

template <typename R, int RB, int H, int L, int PAULI, bool GRAD>
__device__ __forceinline__ void gd_rot2(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t &op,
                                        const cpx<R> *trig, double *gout, bool lane0) {
    if constexpr (GRAD) {
        const double g = wave_sum<64>(gr_rot2<R, RB, H, L>(v, w, PAULI));
        if (lane0) gout[op.grad_slot] = 2.0 * g;
    }
    R cs, sn;
    op_sincos<R>(trig, op, cs, sn);
    ap_rot2<R, RB, H, L>(w, PAULI, cs, sn);
    ap_rot2<R, RB, H, L>(v, PAULI, cs, sn);
}

// The generic kernels' group store: 2·(wave sum) of the first `count` of the
// 2^K gradient terms into the wave's partial row (gout[slot0 + j]).
template <int K>
__device__ __forceinline__ void gd_grad_group(double (&g)[1 << K], double *gout, int slot0, int count) {
    const double s = wave_sum_many<K, 64>(g);
    const int lane = threadIdx.x & 63;
    const int j = rev_low_bits<K>(lane);
    if (lane < (1 << K) && j < count) gout[slot0 + j] = 2.0 * s;
}

}  // namespace svh

#include "svh_patterns.h"

namespace svh {

// Global offset of this thread's register 0 in a layout with run-time tb.
__device__ __forceinline__ uint64_t thr_global_rt(const svh_stage_t &S, int tid, int tb) {
    uint64_t a = 0;
    for (int i = 0; i < tb; ++i) a |= (uint64_t)((tid >> i) & 1) << S.tgbit[i];
    return a;
}

// Tile base address: tile index deposited into the address bits outside the tile.
__device__ __forceinline__ uint64_t tile_base(const svh_pass_t *__restrict__ pass, int n, uint64_t wg) {
    uint64_t tmask = 0;
    for (int i = 0; i < pass->L; ++i) tmask |= 1ull << pass->bits[i];
    uint64_t base = 0, g = wg;
    for (int b = 0; b < n && g; ++b) {
        if (!((tmask >> b) & 1)) {
            base |= (g & 1ull) << b;
            g >>= 1;
        }
    }
    return base;
}

template <typename R, int RB, bool DUAL>
__device__ __forceinline__ bool run_pattern(int pid, cpx<R> (&v)[1 << RB], cpx<R> (&w)[DUAL ? (1 << RB) : 1],
                                            const svh_op_t *ops, const cpx<R> *trig, double *gout, bool lane0) {
    if constexpr (DUAL) {
        if constexpr (RB == 3) return run_pattern3d<R, 3>(pid, v, w, ops, trig, gout, lane0);
        else return run_pattern4d<R, 4>(pid, v, w, ops, trig, gout, lane0);
    } else {
        if constexpr (RB == 3) return run_pattern3f<R, 3>(pid, v, ops, trig);
        else return run_pattern4f<R, 4>(pid, v, ops, trig);
    }
}

// One tile of one pass.  st = the pass's first stage record, ops = the
// program's op array; lds (and, for λ, lds2, or lds + LOFF amplitudes when
// LOFF > 0) hold 2^L amplitudes each.  The tile is loaded from psi and stored
// to out (the same buffer except for the out-of-place last pass of a program
// that ends with a bit permutation).
//
// LDS addressing: a layout puts register j of thread t at slot
// swz(local(t)) ^ lsw[j].  The thread's swizzled base (in bytes) is computed
// once per layout and serves both the exchange that reads a layout and the
// next one that writes it; slot j then costs one XOR with the uniform
// lsw[j]·sizeof(amplitude), and a constant λ offset goes into the LDS
// instruction's offset field.
template <typename R, int RB, bool DUAL, int LOFF = 0>
__device__ __forceinline__ void gtile(const cpx<R> *psi, cpx<R> *out, cpx<R> *__restrict__ lam, int n,
                                      const svh_pass_t *__restrict__ pass, const svh_stage_t *__restrict__ st,
                                      const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig,
                                      double *__restrict__ gpart, uint64_t wg, cpx<R> *lds, cpx<R> *lds2) {
    constexpr int NR = 1 << RB;
    constexpr int ES = (int)sizeof(cpx<R>);
    char *const lb = reinterpret_cast<char *>(lds);
    char *const lb2 = LOFF > 0 ? lb + LOFF * ES : reinterpret_cast<char *>(lds2);
    const int tid = threadIdx.x;
    const int tb = pass->tb;
    const bool active = tid < (1 << tb);
    const int ns = pass->n_stages;
    const uint64_t base = tile_base(pass, n, wg);
    const bool lane0 = (tid & 63) == 0;
    double *gout = nullptr;
    if constexpr (DUAL) gout = gpart + (((wg << (tb - 6)) + (uint64_t)(tid >> 6)) * (uint64_t)pass->n_grad);

    cpx<R> v[NR];
    cpx<R> w[DUAL ? NR : 1];
    if (active) {
        const svh_stage_t &S = st[0];
        const uint64_t t = base | thr_global_rt(S, tid, tb);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            v[j] = psi[t + S.goff[j]];
            if constexpr (DUAL) w[j] = lam[t + S.goff[j]];
        }
    }
    // byte offset of this thread's slot 0 in the current stage's layout
    int tcur = ns > 1 ? lds_swz(thr_local<RB>(st[0], tid)) * ES : 0;
    for (int s = 0; s < ns; ++s) {
        const svh_stage_t &S = st[s];
        if (s > 0) {
            const svh_stage_t &Pv = st[s - 1];
            if (s > 1) __syncthreads();  // the previous exchange's reads are done
            if (active) {
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const int o = tcur ^ (Pv.lsw[j] * ES);
                    *reinterpret_cast<cpx<R> *>(lb + o) = v[j];
                    if constexpr (DUAL) *reinterpret_cast<cpx<R> *>(lb2 + o) = w[j];
                }
            }
            __syncthreads();
            tcur = lds_swz(thr_local<RB>(S, tid)) * ES;
            if (active) {
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const int o = tcur ^ (S.lsw[j] * ES);
                    v[j] = *reinterpret_cast<const cpx<R> *>(lb + o);
                    if constexpr (DUAL) w[j] = *reinterpret_cast<const cpx<R> *>(lb2 + o);
                }
            }
        }
        if (active) {
            const svh_op_t *so = ops + S.op_begin;
            const int pid = S.pattern;
            if (pid < 0 || !run_pattern<R, RB, DUAL>(pid, v, w, so, trig, gout, lane0)) {
                const int nop = S.op_end - S.op_begin;
                for (int o = 0; o < nop; ++o) {
                    double g = 0.0;
                    bool has_g = false;
                    apply_op<R, RB, DUAL>(v, w, so[o], trig, &g, &has_g);
                    if constexpr (DUAL) {
                        if (has_g) {
                            g = wave_sum<64>(g);
                            if (lane0) gout[so[o].grad_slot] = 2.0 * g;
                        }
                    }
                }
            }
        }
    }
    // store through the last layout, or through the remap record after it
    const bool remap = (pass->flags & SVH_PASSF_REMAP_STORE) != 0;
    if (remap && ns == 1) __syncthreads();  // every wave has loaded before any wave stores
    if (active) {
        const svh_stage_t &S = st[remap ? ns : ns - 1];
        const uint64_t t = base | thr_global_rt(S, tid, tb);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            out[t + S.goff[j]] = v[j];
            if constexpr (DUAL) lam[t + S.goff[j]] = w[j];
        }
    }
}

}  // namespace svh
