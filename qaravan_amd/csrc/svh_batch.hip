// Batch helpers: many independent circuits advance together (see
// svh_batch_tile in svh_tile.hip for the fused passes).  Every kernel here maps
// launch blocks to items through a sorted begin table, so one launch serves
// circuits of any size mix.
#include "svh_common.h"

namespace svh {

template <typename T>
__device__ __forceinline__ int find_item(const T *items, int nitems, int64_t b) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].slot_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One block per (item, slot): fixed-order sum over the partial rows.
__global__ void __launch_bounds__(256) k_batch_reduce(const svh_batch_red_t *__restrict__ items, int nitems) {
    const int64_t b = blockIdx.x;
    const int i = find_item(items, nitems, b);
    const svh_batch_red_t &it = items[i];
    const int slot = (int)(b - it.slot_begin);
    const double *part = reinterpret_cast<const double *>(it.part);
    double acc = 0.0;
    for (int r = threadIdx.x; r < it.nblocks; r += blockDim.x) acc += part[(int64_t)r * it.ng + slot];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int32_t *sp = reinterpret_cast<const int32_t *>(it.slot_param);
        double *out = reinterpret_cast<double *>(it.out);
        atomicAdd(&out[sp[slot]], red[0]);
        if (it.tstamp) {
            __threadfence();
            atomicMax(reinterpret_cast<unsigned long long *>(it.tstamp),
                      (unsigned long long)wall_clock64());
        }
    }
}

struct HamTerm {
    uint64_t x, z;
    double cre, cim;
};

__device__ __forceinline__ double psign(uint64_t v) { return (__popcll(v) & 1) ? -1.0 : 1.0; }

__device__ __forceinline__ int find_ham(const svh_batch_ham_t *items, int nitems, int64_t b) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// 256 threads x 4 amplitudes per block; λ_i = Σ_t c_t i^{ny} (-1)^{pc((i^x)&z)} ψ[i^x]
template <typename R>
__global__ void __launch_bounds__(256) k_batch_ham(const svh_batch_ham_t *__restrict__ items, int nitems) {
    const int64_t b = blockIdx.x;
    const svh_batch_ham_t &it = items[find_ham(items, nitems, b)];
    // Table addresses converted through the global address space: the uniform
    // term records then load through the scalar unit.
    const cpx<R> *psi = (const cpx<R> *)(__attribute__((address_space(1))) const cpx<R> *)it.psi;
    cpx<R> *lam = (cpx<R> *)(__attribute__((address_space(1))) cpx<R> *)it.lam;
    const HamTerm *terms = (const HamTerm *)(__attribute__((address_space(1))) const HamTerm *)it.terms;
    const uint64_t i0 = (uint64_t)(b - it.block_begin) * 1024;
    // terms outside, the thread's 4 amplitudes inside: 4 independent gathers
    // per term in flight (each amplitude still sums its terms in table order)
    double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
    int t = 0;
    for (; t + 2 <= it.nterms; t += 2) {  // 2 terms x 4 amplitudes: 8 gathers in flight
        const HamTerm T[2] = {terms[t], terms[t + 1]};
        cpx<R> v[2][4];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[u][k] = psi[(i0 + (uint64_t)k * 256 + threadIdx.x) ^ T[u].x];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t j = (i0 + (uint64_t)k * 256 + threadIdx.x) ^ T[u].x;
                const double s = psign(j & T[u].z);
                ar[k] += s * (T[u].cre * v[u][k].x - T[u].cim * v[u][k].y);
                ai[k] += s * (T[u].cre * v[u][k].y + T[u].cim * v[u][k].x);
            }
    }
    for (; t < it.nterms; ++t) {
        const HamTerm T = terms[t];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t j = (i0 + (uint64_t)k * 256 + threadIdx.x) ^ T.x;
            const cpx<R> v = psi[j];
            const double s = psign(j & T.z);
            ar[k] += s * (T.cre * v.x - T.cim * v.y);
            ai[k] += s * (T.cre * v.y + T.cim * v.x);
        }
    }
    double e = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t i = i0 + (uint64_t)k * 256 + threadIdx.x;
        lam[i] = cpx<R>{(R)ar[k], (R)ai[k]};
        const cpx<R> p = psi[i];
        e += (double)p.x * ar[k] + (double)p.y * ai[k];  // Re(conj(ψ_i) λ_i)
    }
    __shared__ double red[256];
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) reinterpret_cast<double *>(it.epart)[b - it.block_begin] = red[0];
}

template <typename R>
__global__ void k_batch_init(cpx<R> *__restrict__ buf, int64_t total, const int64_t *__restrict__ offs, int no) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        int lo = 0, hi = no - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        buf[i] = cpx<R>{(R)(offs[lo] == i ? 1 : 0), (R)0};
    }
}

template <typename R>
__global__ void k_batch_trig(const double *__restrict__ theta, int64_t np, cpx<R> *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) {
        double sn, cs;
        sincos(theta[i], &sn, &cs);
        out[i] = cpx<R>{(R)cs, (R)sn};
    }
}

__global__ void k_timestamp(uint64_t *out) {
    if (threadIdx.x == 0) out[0] = wall_clock64();
}

}  // namespace svh

using namespace svh;

extern "C" {

int svh_batch_reduce(const void *red_dev, int nitems, int64_t total_blocks, void *stream) {
    if (!red_dev || nitems <= 0) return fail(-1, "empty reduction table");
    if (total_blocks <= 0 || total_blocks > 0x7fffffff) return fail(-4, "bad block count");
    k_batch_reduce<<<(unsigned)total_blocks, 256, 0, (hipStream_t)stream>>>(
        static_cast<const svh_batch_red_t *>(red_dev), nitems);
    return checked_launch("k_batch_reduce");
}

int svh_batch_hamiltonian(const void *ham_dev, int nitems, int64_t total_blocks, int dtype, void *stream) {
    if (!ham_dev || nitems <= 0) return fail(-1, "empty Hamiltonian table");
    if (total_blocks <= 0 || total_blocks > 0x7fffffff) return fail(-4, "bad block count");
    const auto *it = static_cast<const svh_batch_ham_t *>(ham_dev);
    if (dtype == SVH_C64)
        k_batch_ham<float><<<(unsigned)total_blocks, 256, 0, (hipStream_t)stream>>>(it, nitems);
    else if (dtype == SVH_C128)
        k_batch_ham<double><<<(unsigned)total_blocks, 256, 0, (hipStream_t)stream>>>(it, nitems);
    else
        return fail(-3, "bad dtype");
    return checked_launch("k_batch_ham");
}

int svh_batch_init(void *buf, int64_t total, const int64_t *offsets_dev, int noffsets, int dtype, void *stream) {
    if (!buf || !offsets_dev || noffsets <= 0 || total <= 0) return fail(-1, "bad batch init arguments");
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (dtype == SVH_C64)
        k_batch_init<float><<<(unsigned)g, 256, 0, (hipStream_t)stream>>>((cpx<float> *)buf, total, offsets_dev,
                                                                           noffsets);
    else if (dtype == SVH_C128)
        k_batch_init<double><<<(unsigned)g, 256, 0, (hipStream_t)stream>>>((cpx<double> *)buf, total, offsets_dev,
                                                                            noffsets);
    else
        return fail(-3, "bad dtype");
    return checked_launch("k_batch_init");
}

int svh_batch_trig(const double *theta_dev, int64_t np, void *trig_dev, int dtype, void *stream) {
    if (!theta_dev || !trig_dev || np <= 0) return fail(-1, "bad trig arguments");
    const unsigned g = (unsigned)((np + 255) / 256);
    if (dtype == SVH_C64)
        k_batch_trig<float><<<g, 256, 0, (hipStream_t)stream>>>(theta_dev, np, (cpx<float> *)trig_dev);
    else if (dtype == SVH_C128)
        k_batch_trig<double><<<g, 256, 0, (hipStream_t)stream>>>(theta_dev, np, (cpx<double> *)trig_dev);
    else
        return fail(-3, "bad dtype");
    return checked_launch("k_batch_trig");
}

int svh_timestamp(uint64_t *out_dev, void *stream) {
    if (!out_dev) return fail(-1, "NULL pointer");
    k_timestamp<<<1, 64, 0, (hipStream_t)stream>>>(out_dev);
    return checked_launch("k_timestamp");
}

}  // extern "C"
