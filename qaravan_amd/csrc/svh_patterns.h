// Generated by `python -m qaravan_amd.patterns` (do not edit): straight-line
// code for the stage op sequences of the pattern library (svh_patterns.json).
#pragma once
#include "svh_gtile.h"

namespace svh {

template <typename R, int RB>
__device__ __forceinline__ void pat3f_0(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_1(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[1], trig);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_2(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_3(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_4(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[9], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_5(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[1], trig);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[3], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[4], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[5], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[6], trig);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_6(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_7(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[13], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[14], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_8(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[8], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_9(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[10], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_10(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[10], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[12], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[13], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_11(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    ap_cz<R, RB, 2, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_12(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    ap_cz<R, RB, 2, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[12], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[13], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_13(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_14(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[12], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_15(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[11], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[12], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_16(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[12], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[15], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[16], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[17], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[18], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_17(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[11], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[12], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[15], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[16], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[17], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[18], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_18(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[12], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[15], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[16], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[17], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[18], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[20], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[21], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_19(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[11], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[12], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[15], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[16], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[17], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[18], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[20], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[21], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_20(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[3], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_21(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[3], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_22(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_23(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[3], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[14], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[15], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[16], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[17], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_24(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[3], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[4], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[5], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[6], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[7], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_25(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[3], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[14], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[15], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[16], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[17], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[19], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[20], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_26(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[3], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[4], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[5], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[6], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[7], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[10], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[11], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_27(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_28(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[9], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[10], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[11], trig);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[13], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[14], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_29(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_30(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[1], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_31(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[2], trig);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_32(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[1], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_33(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[2], trig);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_34(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[1], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[7], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_35(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    ap_cz<R, RB, 2, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_36(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[2], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_37(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    ap_cz<R, RB, 2, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[4], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[5], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_38(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_39(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_40(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[7], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_41(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[10], trig);
    ap_cz<R, RB, 2, 1>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_42(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    sp_rot1<R, 0, 2, RB>(v, ops[0], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[2], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[3], trig);
    ap_cz<R, RB, 2, 0>(v);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 0, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[8], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[9], trig);
    ap_cz<R, RB, 2, 0>(v);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_43(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[6], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[7], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_44(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[8], trig);
}
template <typename R, int RB>
__device__ __forceinline__ void pat3f_45(cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    ap_cz<R, RB, 1, 0>(v);
    sp_rot1<R, 0, 2, RB>(v, ops[1], trig);
    sp_rot1<R, 1, 2, RB>(v, ops[2], trig);
    sp_rot1<R, 0, 3, RB>(v, ops[3], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[4], trig);
    ap_cz<R, RB, 1, 0>(v);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 1, 2, RB>(v, ops[7], trig);
    sp_rot1<R, 2, 2, RB>(v, ops[8], trig);
    sp_rot1<R, 1, 3, RB>(v, ops[9], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[10], trig);
    ap_cz<R, RB, 2, 1>(v);
    sp_rot1<R, 2, 2, RB>(v, ops[12], trig);
    sp_rot1<R, 2, 3, RB>(v, ops[13], trig);
}
// pattern 3f: returns false when pid is not in the library
template <typename R, int RB>
__device__ __forceinline__ bool run_pattern3f(int pid, cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    switch (pid) {
        case 0: pat3f_0<R, RB>(v, ops, trig); return true;
        case 1: pat3f_1<R, RB>(v, ops, trig); return true;
        case 2: pat3f_2<R, RB>(v, ops, trig); return true;
        case 3: pat3f_3<R, RB>(v, ops, trig); return true;
        case 4: pat3f_4<R, RB>(v, ops, trig); return true;
        case 5: pat3f_5<R, RB>(v, ops, trig); return true;
        case 6: pat3f_6<R, RB>(v, ops, trig); return true;
        case 7: pat3f_7<R, RB>(v, ops, trig); return true;
        case 8: pat3f_8<R, RB>(v, ops, trig); return true;
        case 9: pat3f_9<R, RB>(v, ops, trig); return true;
        case 10: pat3f_10<R, RB>(v, ops, trig); return true;
        case 11: pat3f_11<R, RB>(v, ops, trig); return true;
        case 12: pat3f_12<R, RB>(v, ops, trig); return true;
        case 13: pat3f_13<R, RB>(v, ops, trig); return true;
        case 14: pat3f_14<R, RB>(v, ops, trig); return true;
        case 15: pat3f_15<R, RB>(v, ops, trig); return true;
        case 16: pat3f_16<R, RB>(v, ops, trig); return true;
        case 17: pat3f_17<R, RB>(v, ops, trig); return true;
        case 18: pat3f_18<R, RB>(v, ops, trig); return true;
        case 19: pat3f_19<R, RB>(v, ops, trig); return true;
        case 20: pat3f_20<R, RB>(v, ops, trig); return true;
        case 21: pat3f_21<R, RB>(v, ops, trig); return true;
        case 22: pat3f_22<R, RB>(v, ops, trig); return true;
        case 23: pat3f_23<R, RB>(v, ops, trig); return true;
        case 24: pat3f_24<R, RB>(v, ops, trig); return true;
        case 25: pat3f_25<R, RB>(v, ops, trig); return true;
        case 26: pat3f_26<R, RB>(v, ops, trig); return true;
        case 27: pat3f_27<R, RB>(v, ops, trig); return true;
        case 28: pat3f_28<R, RB>(v, ops, trig); return true;
        case 29: pat3f_29<R, RB>(v, ops, trig); return true;
        case 30: pat3f_30<R, RB>(v, ops, trig); return true;
        case 31: pat3f_31<R, RB>(v, ops, trig); return true;
        case 32: pat3f_32<R, RB>(v, ops, trig); return true;
        case 33: pat3f_33<R, RB>(v, ops, trig); return true;
        case 34: pat3f_34<R, RB>(v, ops, trig); return true;
        case 35: pat3f_35<R, RB>(v, ops, trig); return true;
        case 36: pat3f_36<R, RB>(v, ops, trig); return true;
        case 37: pat3f_37<R, RB>(v, ops, trig); return true;
        case 38: pat3f_38<R, RB>(v, ops, trig); return true;
        case 39: pat3f_39<R, RB>(v, ops, trig); return true;
        case 40: pat3f_40<R, RB>(v, ops, trig); return true;
        case 41: pat3f_41<R, RB>(v, ops, trig); return true;
        case 42: pat3f_42<R, RB>(v, ops, trig); return true;
        case 43: pat3f_43<R, RB>(v, ops, trig); return true;
        case 44: pat3f_44<R, RB>(v, ops, trig); return true;
        case 45: pat3f_45<R, RB>(v, ops, trig); return true;
        default: return false;
    }
}

constexpr int kPatCount3f = 46;
constexpr int kPatOff3f[] = {0, 8, 11, 12, 25, 35, 43, 50, 65, 74, 86, 100, 112, 126, 133, 148, 163, 183, 203, 225, 247, 253, 267, 269, 288, 298, 319, 331, 344, 359, 366, 369, 373, 379, 385, 393, 397, 400, 406, 409, 415, 423, 435, 446, 454, 463, 477};
constexpr signed char kPatOps3f[][5] = {{5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {4, 2, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 2, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 2, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 2, 0, 2, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 2, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 2, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 0, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 2, 0}, {5, 0, 0, 2, 0}, {5, 2, 0, 3, 0}, {5, 0, 0, 3, 0}, {4, 2, 0, 0, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 0}, {5, 1, 0, 2, 0}, {5, 0, 0, 3, 0}, {5, 1, 0, 3, 0}, {4, 1, 0, 0, 0}, {4, 2, 1, 0, 0}, {5, 1, 0, 2, 0}, {5, 2, 0, 2, 0}, {5, 1, 0, 3, 0}, {5, 2, 0, 3, 0}, {4, 2, 1, 0, 0}, {5, 2, 0, 2, 0}, {5, 2, 0, 3, 0}};

template <typename R, int RB>
__device__ __forceinline__ void pat3d_0(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq2, gq3, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[2].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_1(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[6], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_2(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq2, gq3, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[2].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_3(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq2, gq3, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[2].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_4(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    { double gv[4] = {gq1, gq2, gq3, 0.0}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 3); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_5(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_6(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[6], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, gq8, gq9, gq10}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_7(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[5], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq8, gq9, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_8(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_9(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq3, gq6, gq7, gq8, gq9, gq10}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_10(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    { double gv[2] = {gq1, gq2}; gd_grad_group<1>(gv, gout, ops[1].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_11(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    { double gv[4] = {gq2, gq3, gq4, 0.0}; gd_grad_group<2>(gv, gout, ops[2].grad_slot, 3); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_12(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[6], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, gq8, gq9, gq10}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
    const double gq11 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[1] = {gq11}; gd_grad_group<0>(gv, gout, ops[11].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_13(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[5], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq8, gq9, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[1] = {gq12}; gd_grad_group<0>(gv, gout, ops[12].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_14(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq3, gq6, gq7, gq8, gq9, gq10}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[16], trig, gout, lane0);
    const double gq17 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[17], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq17, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_15(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_16(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq7, gq8, gq9, gq10, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 7); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_17(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq3, gq6, gq7, gq8, gq9, gq10}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[16], trig, gout, lane0);
    const double gq17 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[17], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq19 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[19], trig, gout, lane0);
    const double gq20 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[20], trig, gout, lane0);
    const double gq21 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[21], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq17, gq19, gq20, gq21}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_18(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq10 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[1] = {gq12}; gd_grad_group<0>(gv, gout, ops[12].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_19(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq7, gq8, gq9, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_20(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_21(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[3], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[10], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq3, gq6, gq7, gq8, gq9, gq10}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[16], trig, gout, lane0);
    const double gq17 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[17], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq19 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[19], trig, gout, lane0);
    const double gq20 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[20], trig, gout, lane0);
    const double gq21 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[21], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq17, gq19, gq20, gq21}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 8); }
    const double gq22 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[22], trig, gout, lane0);
    { double gv[1] = {gq22}; gd_grad_group<0>(gv, gout, ops[22].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_22(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq10 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[13], trig, gout, lane0);
    { double gv[2] = {gq12, gq13}; gd_grad_group<1>(gv, gout, ops[12].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_23(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_24(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_25(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq10 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_26(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq10 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[1] = {gq12}; gd_grad_group<0>(gv, gout, ops[12].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_27(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq10 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[13], trig, gout, lane0);
    { double gv[2] = {gq12, gq13}; gd_grad_group<1>(gv, gout, ops[12].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_28(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[9], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 7); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_29(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[9], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq12 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, gq12}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq13 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[16], trig, gout, lane0);
    { double gv[4] = {gq13, gq14, gq15, gq16}; gd_grad_group<2>(gv, gout, ops[13].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_30(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[9], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq12 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, gq12}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq13 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[16], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq18 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[18], trig, gout, lane0);
    const double gq19 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[19], trig, gout, lane0);
    const double gq20 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[20], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq18, gq19, gq20, 0.0}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 7); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_31(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[9], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq12 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, gq12}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq13 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[16], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq18 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[18], trig, gout, lane0);
    const double gq19 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[19], trig, gout, lane0);
    const double gq20 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[20], trig, gout, lane0);
    const double gq21 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[21], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq18, gq19, gq20, gq21}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_32(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[9], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 7); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_33(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[9], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, gq12}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq13 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[16], trig, gout, lane0);
    { double gv[4] = {gq13, gq14, gq15, gq16}; gd_grad_group<2>(gv, gout, ops[13].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_34(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[9], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, gq12}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq13 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[16], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq18 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[18], trig, gout, lane0);
    const double gq19 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[19], trig, gout, lane0);
    const double gq20 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[20], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq18, gq19, gq20, 0.0}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 7); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_35(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[8], trig, gout, lane0);
    const double gq9 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[9], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq9, gq12}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq13 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[13], trig, gout, lane0);
    const double gq14 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[14], trig, gout, lane0);
    const double gq15 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[16], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq18 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[18], trig, gout, lane0);
    const double gq19 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[19], trig, gout, lane0);
    const double gq20 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[20], trig, gout, lane0);
    const double gq21 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[21], trig, gout, lane0);
    { double gv[8] = {gq13, gq14, gq15, gq16, gq18, gq19, gq20, gq21}; gd_grad_group<3>(gv, gout, ops[13].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_36(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[6], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_37(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[6], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, gq9, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[13], trig, gout, lane0);
    { double gv[2] = {gq12, gq13}; gd_grad_group<1>(gv, gout, ops[12].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_38(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[6], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, gq9, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[13], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq15 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[16], trig, gout, lane0);
    const double gq17 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[17], trig, gout, lane0);
    { double gv[8] = {gq12, gq13, gq15, gq16, gq17, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[12].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_39(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[6], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq9 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[9], trig, gout, lane0);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq5, gq6, gq9, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[13], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq15 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[15], trig, gout, lane0);
    const double gq16 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[16], trig, gout, lane0);
    const double gq17 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[17], trig, gout, lane0);
    const double gq18 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[18], trig, gout, lane0);
    { double gv[8] = {gq12, gq13, gq15, gq16, gq17, gq18, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[12].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_40(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    { double gv[2] = {gq1, gq2}; gd_grad_group<1>(gv, gout, ops[1].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_41(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq1, gq2, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_42(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[6], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq4, gq5, gq6, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_43(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    { double gv[1] = {gq1}; gd_grad_group<0>(gv, gout, ops[1].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_44(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    { double gv[4] = {gq1, gq3, gq4, 0.0}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 3); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_45(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq1, gq3, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_46(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    ap_cz<R, RB, 2, 1>(v); ap_cz<R, RB, 2, 1>(w);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[3], trig, gout, lane0);
    { double gv[2] = {gq1, gq3}; gd_grad_group<1>(gv, gout, ops[1].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_47(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_48(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_49(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_50(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq10 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[1] = {gq12}; gd_grad_group<0>(gv, gout, ops[12].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_51(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq10 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[13], trig, gout, lane0);
    { double gv[2] = {gq12, gq13}; gd_grad_group<1>(gv, gout, ops[12].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_52(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[6], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq4, gq5, gq6, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_53(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq4, gq5, gq6, gq7, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_54(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[6], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq4, gq5, gq6, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_55(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[7], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq4, gq5, gq6, gq7, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_56(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq1, gq2, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_57(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[12], trig, gout, lane0);
    { double gv[1] = {gq12}; gd_grad_group<0>(gv, gout, ops[12].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_58(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[8], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq10 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[10], trig, gout, lane0);
    const double gq11 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[11], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, gq10, gq11}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 8); }
    const double gq12 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[12], trig, gout, lane0);
    const double gq13 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[13], trig, gout, lane0);
    { double gv[2] = {gq12, gq13}; gd_grad_group<1>(gv, gout, ops[12].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_59(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[1], trig, gout, lane0);
    { double gv[1] = {gq1}; gd_grad_group<0>(gv, gout, ops[1].grad_slot, 1); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_60(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[1], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    { double gv[4] = {gq1, gq3, gq4, 0.0}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 3); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_61(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[1], trig, gout, lane0);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq3 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 2>(v, w, 2);
    gd_rot1<R, RB, 2, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[4] = {gq1, gq3, gq4, gq5}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_62(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[7], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_63(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[2], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    const double gq6 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[7], trig, gout, lane0);
    const double gq8 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[8], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq5, gq6, gq7, gq8, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 6); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_64(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[2], trig, gout, lane0);
    { double gv[2] = {gq1, gq2}; gd_grad_group<1>(gv, gout, ops[1].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_65(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    { double gv[2] = {gq1, gq2}; gd_grad_group<1>(gv, gout, ops[1].grad_slot, 2); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_66(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[4], trig, gout, lane0);
    const double gq5 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[5], trig, gout, lane0);
    { double gv[8] = {gq1, gq2, gq3, gq4, gq5, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[1].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_67(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[4], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq6, gq7, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 5); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_68(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq1 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[1], trig, gout, lane0);
    const double gq2 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[4], trig, gout, lane0);
    { double gv[4] = {gq1, gq2, gq3, gq4}; gd_grad_group<2>(gv, gout, ops[1].grad_slot, 4); }
}
template <typename R, int RB>
__device__ __forceinline__ void pat3d_69(cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    ap_cz<R, RB, 2, 0>(v); ap_cz<R, RB, 2, 0>(w);
    const double gq2 = gr_rot1<R, RB, 2>(v, w, 3);
    gd_rot1<R, RB, 2, 3, false>(v, w, ops[2], trig, gout, lane0);
    const double gq3 = gr_rot1<R, RB, 1>(v, w, 2);
    gd_rot1<R, RB, 1, 2, false>(v, w, ops[3], trig, gout, lane0);
    const double gq4 = gr_rot1<R, RB, 0>(v, w, 2);
    gd_rot1<R, RB, 0, 2, false>(v, w, ops[4], trig, gout, lane0);
    ap_cz<R, RB, 1, 0>(v); ap_cz<R, RB, 1, 0>(w);
    const double gq6 = gr_rot1<R, RB, 1>(v, w, 3);
    gd_rot1<R, RB, 1, 3, false>(v, w, ops[6], trig, gout, lane0);
    const double gq7 = gr_rot1<R, RB, 0>(v, w, 3);
    gd_rot1<R, RB, 0, 3, false>(v, w, ops[7], trig, gout, lane0);
    { double gv[8] = {gq2, gq3, gq4, gq6, gq7, 0.0, 0.0, 0.0}; gd_grad_group<3>(gv, gout, ops[2].grad_slot, 5); }
}
// pattern 3d: returns false when pid is not in the library
template <typename R, int RB>
__device__ __forceinline__ bool run_pattern3d(int pid, cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    switch (pid) {
        case 0: pat3d_0<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 1: pat3d_1<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 2: pat3d_2<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 3: pat3d_3<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 4: pat3d_4<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 5: pat3d_5<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 6: pat3d_6<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 7: pat3d_7<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 8: pat3d_8<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 9: pat3d_9<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 10: pat3d_10<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 11: pat3d_11<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 12: pat3d_12<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 13: pat3d_13<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 14: pat3d_14<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 15: pat3d_15<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 16: pat3d_16<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 17: pat3d_17<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 18: pat3d_18<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 19: pat3d_19<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 20: pat3d_20<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 21: pat3d_21<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 22: pat3d_22<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 23: pat3d_23<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 24: pat3d_24<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 25: pat3d_25<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 26: pat3d_26<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 27: pat3d_27<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 28: pat3d_28<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 29: pat3d_29<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 30: pat3d_30<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 31: pat3d_31<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 32: pat3d_32<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 33: pat3d_33<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 34: pat3d_34<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 35: pat3d_35<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 36: pat3d_36<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 37: pat3d_37<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 38: pat3d_38<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 39: pat3d_39<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 40: pat3d_40<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 41: pat3d_41<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 42: pat3d_42<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 43: pat3d_43<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 44: pat3d_44<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 45: pat3d_45<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 46: pat3d_46<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 47: pat3d_47<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 48: pat3d_48<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 49: pat3d_49<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 50: pat3d_50<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 51: pat3d_51<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 52: pat3d_52<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 53: pat3d_53<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 54: pat3d_54<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 55: pat3d_55<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 56: pat3d_56<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 57: pat3d_57<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 58: pat3d_58<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 59: pat3d_59<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 60: pat3d_60<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 61: pat3d_61<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 62: pat3d_62<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 63: pat3d_63<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 64: pat3d_64<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 65: pat3d_65<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 66: pat3d_66<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 67: pat3d_67<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 68: pat3d_68<R, RB>(v, w, ops, trig, gout, lane0); return true;
        case 69: pat3d_69<R, RB>(v, w, ops, trig, gout, lane0); return true;
        default: return false;
    }
}

constexpr int kPatCount3d = 70;
constexpr int kPatOff3d[] = {0, 6, 13, 19, 25, 29, 36, 47, 59, 68, 79, 82, 87, 99, 112, 130, 139, 150, 172, 185, 197, 205, 228, 242, 251, 260, 272, 285, 299, 309, 326, 347, 369, 379, 396, 417, 439, 446, 460, 478, 497, 500, 506, 513, 515, 520, 526, 530, 539, 551, 560, 573, 587, 594, 602, 609, 617, 623, 636, 650, 652, 657, 663, 671, 680, 683, 686, 692, 700, 705, 713};
constexpr signed char kPatOps3d[][5] = {{4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 2, 1, 0, 0}, {4, 1, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 1, 0, 3, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 1, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 1, 0, 3, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 1, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {4, 2, 1, 0, 0}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 2, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 1}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 2, 1}, {5, 2, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 2, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}, {4, 1, 0, 0, 0}, {5, 0, 0, 3, 1}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 1, 0, 3, 1}, {4, 1, 0, 0, 0}, {4, 2, 0, 0, 0}, {5, 2, 0, 3, 1}, {5, 1, 0, 2, 1}, {5, 0, 0, 2, 1}, {4, 1, 0, 0, 0}, {5, 1, 0, 3, 1}, {5, 0, 0, 3, 1}};

// pattern 4f: returns false when pid is not in the library
template <typename R, int RB>
__device__ __forceinline__ bool run_pattern4f(int pid, cpx<R> (&v)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig) {
    switch (pid) {
        default: return false;
    }
}

constexpr int kPatCount4f = 0;
constexpr int kPatOff4f[] = {0};
constexpr signed char kPatOps4f[][5] = {{0, 0, 0, 0, 0}};

// pattern 4d: returns false when pid is not in the library
template <typename R, int RB>
__device__ __forceinline__ bool run_pattern4d(int pid, cpx<R> (&v)[1 << RB], cpx<R> (&w)[1 << RB], const svh_op_t *__restrict__ ops, const cpx<R> *__restrict__ trig, double *__restrict__ gout, bool lane0) {
    switch (pid) {
        default: return false;
    }
}

constexpr int kPatCount4d = 0;
constexpr int kPatOff4d[] = {0};
constexpr signed char kPatOps4d[][5] = {{0, 0, 0, 0, 0}};

}  // namespace svh
