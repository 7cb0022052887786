// Shared device/host helpers for the gfx950 statevector kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>
#include <type_traits>

#include "../../include/qaravan_hip.h"

namespace svh {

// Interleaved complex amplitude; alignment makes every load/store one
// 8-byte (float) or 16-byte (double) vector access.
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ R ffma(R a, R b, R c) {
    if constexpr (std::is_same<R, float>::value) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}
// a*b with explicit FMAs: 2 VALU ops per component
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
    return {ffma(a.x, b.x, -a.y * b.y), ffma(a.x, b.y, a.y * b.x)};
}
template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) {
    return {a.x + b.x, a.y + b.y};
}
// a*b + c: 2 FMAs per component
template <typename R>
__device__ __forceinline__ cpx<R> cfma(cpx<R> a, cpx<R> b, cpx<R> c) {
    return {ffma(a.x, b.x, ffma(-a.y, b.y, c.x)), ffma(a.x, b.y, ffma(a.y, b.x, c.y))};
}
template <typename R>
__device__ __forceinline__ cpx<R> cneg(cpx<R> a) {
    return {-a.x, -a.y};
}
// acc + Im(conj(l) * z), two fused multiply-adds in the state precision
template <typename R>
__device__ __forceinline__ R im_conj_mul_acc(R acc, cpx<R> l, cpx<R> z) {
    return fma(-l.y, z.x, fma(l.x, z.y, acc));
}

// Insert a zero bit at position B of p.
template <int B>
__device__ __forceinline__ constexpr int ins0(int p) {
    return ((p >> B) << (B + 1)) | (p & ((1 << B) - 1));
}

// Status plumbing (thread-local last error string).
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_status(hipError_t e, const char *where);

inline int checked_launch(const char *where) {
    return hip_status(hipGetLastError(), where);
}

// Raise a kernel's dynamic-LDS limit to `bytes` on the current device, once
// per (device, kernel, bytes); thread-safe (a mutex-guarded set).
int ensure_lds_limit(const void *kern, int bytes, const char *where);
// Resident workgroups per CU of `kern` at (threads, lds) and the CU count of
// the current device; computed once per (device, kernel, threads, lds) under
// a mutex, so concurrent callers on several devices each get their own.
int resident_shape(const void *kern, int threads, size_t lds, int *per_cu, int *cus);

constexpr int kMaxQubits = 36;
constexpr size_t kReduceBlocks = 1024;  // partials per FP64 reduction

}  // namespace svh
