// Unfused kernels: basis init, one dense k-qubit gate, FP64 reductions
// (norm, inner product, Pauli-sum expectation) and H·ψ for Pauli sums.
#include "svh_common.h"

#include <map>
#include <mutex>
#include <set>
#include <tuple>

namespace svh {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_status(hipError_t e, const char *where) {
    if (e == hipSuccess) return 0;
    g_last_error = std::string(where) + ": " + hipGetErrorString(e);
    return (int)e;
}

static std::mutex g_launch_mu;

int ensure_lds_limit(const void *kern, int bytes, const char *where) {
    static std::set<std::tuple<int, const void *, int>> done;
    int dev = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    std::lock_guard<std::mutex> lock(g_launch_mu);
    const auto key = std::make_tuple(dev, kern, bytes);
    if (done.count(key)) return 0;
    if (int e = hip_status(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, bytes), where))
        return e;
    done.insert(key);
    return 0;
}

int resident_shape(const void *kern, int threads, size_t lds, int *per_cu, int *cus) {
    static std::map<std::tuple<int, const void *, int, size_t>, std::pair<int, int>> cache;
    int dev = 0;
    if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
    std::lock_guard<std::mutex> lock(g_launch_mu);
    const auto key = std::make_tuple(dev, kern, threads, lds);
    auto it = cache.find(key);
    if (it == cache.end()) {
        int nb = 0, nc = 0;
        if (int e = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, lds),
                               "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
            return e;
        if (int e = hip_status(hipDeviceGetAttribute(&nc, hipDeviceAttributeMultiprocessorCount, dev),
                               "hipDeviceGetAttribute"))
            return e;
        it = cache.emplace(key, std::make_pair(nb > 0 ? nb : 1, nc > 0 ? nc : 1)).first;
    }
    *per_cu = it->second.first;
    *cus = it->second.second;
    return 0;
}

// ---------------------------------------------------------------- init
template <typename R>
__global__ void k_init_basis(cpx<R> *__restrict__ psi, uint64_t N, uint64_t idx) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        cpx<R> z{(R)(i == idx ? 1 : 0), (R)0};
        psi[i] = z;
    }
}

static unsigned grid_for(uint64_t work, unsigned block, unsigned cap = 65536) {
    uint64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// ---------------------------------------------------------------- dense k-qubit gate
// One thread per amplitude.  A workgroup holds whole groups (the 2^k amplitudes
// that differ only in the gate bits): each thread stages its amplitude in LDS,
// then computes one output row from the group, so the update is in place.
template <typename R>
__global__ void __launch_bounds__(256)
k_dense(cpx<R> *__restrict__ psi, int n, int k, uint64_t gate_mask, int4 b03, int2 b45,
        const cpx<R> *__restrict__ mat, uint64_t N) {
    extern __shared__ unsigned char smem_raw[];
    const int K = 1 << k;
    cpx<R> *msh = reinterpret_cast<cpx<R> *>(smem_raw);  // K*K
    cpx<R> *grp = msh + K * K;                             // 256
    for (int i = threadIdx.x; i < K * K; i += blockDim.x) msh[i] = mat[i];
    const int bits[6] = {b03.x, b03.y, b03.z, b03.w, b45.x, b45.y};

    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = idx < N;
    const int j = (int)(idx & (uint64_t)(K - 1));  // row inside the group
    uint64_t g = idx >> k;                         // group id
    // deposit g into the non-gate bits, j into the gate bits
    uint64_t addr = 0;
    for (int b = 0; b < n; ++b) {
        if (!((gate_mask >> b) & 1)) {
            addr |= (g & 1ull) << b;
            g >>= 1;
        }
    }
    for (int a = 0; a < k; ++a) addr |= (uint64_t)((j >> (k - 1 - a)) & 1) << bits[a];
    if (live) grp[threadIdx.x] = psi[addr];
    __syncthreads();
    if (live) {
        const int g0 = threadIdx.x - j;
        cpx<R> acc{(R)0, (R)0};
        for (int c = 0; c < K; ++c) acc = cfma(msh[j * K + c], grp[g0 + c], acc);
        psi[addr] = acc;
    }
}

// ---------------------------------------------------------------- reductions
// Deterministic two-stage FP64 sums: each block writes one partial (fixed
// tree order), then a single block adds the partials in a fixed order.
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *__restrict__ out) {
    __shared__ double red[NV][256];
#pragma unroll
    for (int c = 0; c < NV; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int c = 0; c < NV; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c * gridDim.x + blockIdx.x] = red[c][0];
    }
}

template <int NV>
__global__ void __launch_bounds__(256) k_final_sum(const double *__restrict__ part, int nblocks,
                                                   double *__restrict__ out) {
    double v[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
#pragma unroll
        for (int c = 0; c < NV; ++c) v[c] += part[c * nblocks + i];
    }
    __shared__ double red[NV][256];
#pragma unroll
    for (int c = 0; c < NV; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int c = 0; c < NV; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = red[c][0];
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_norm2(const cpx<R> *__restrict__ psi, uint64_t N,
                                               double *__restrict__ part) {
    double v[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const cpx<R> a = psi[i];
        v[0] += (double)a.x * a.x + (double)a.y * a.y;
    }
    block_reduce_store<1>(v, part);
}

template <typename R>
__global__ void __launch_bounds__(256) k_inner(const cpx<R> *__restrict__ a,
                                               const cpx<R> *__restrict__ b, uint64_t N,
                                               double *__restrict__ part) {
    double v[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const cpx<R> x = a[i], y = b[i];
        v[0] += (double)x.x * y.x + (double)x.y * y.y;  // Re(conj(x) y)
        v[1] += (double)x.x * y.y - (double)x.y * y.x;  // Im(conj(x) y)
    }
    block_reduce_store<2>(v, part);
}

template <typename R>
__global__ void k_scale(cpx<R> *__restrict__ psi, uint64_t N, const double *__restrict__ norm2) {
    const R s = (R)(1.0 / sqrt(norm2[0]));
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        cpx<R> a = psi[i];
        a.x *= s;
        a.y *= s;
        psi[i] = a;
    }
}

// Pauli term table in device memory: xmask, zmask, coeff, phase class
// (ny mod 4 folded into coefficient as c·i^{ny}: stored as (re, im)).
struct PauliTerm {
    uint64_t x, z;
    double cre, cim;
};

__device__ __forceinline__ double parity_sign(uint64_t v) {
    return (__popcll(v) & 1) ? -1.0 : 1.0;
}

// Σ_t c_t i^{ny} Σ_i conj(ψ[i^x]) (-1)^{pc(i&z)} ψ[i]; real part accumulated.
template <typename R>
__global__ void __launch_bounds__(256) k_pauli_expect(const cpx<R> *__restrict__ psi, uint64_t N,
                                                      const PauliTerm *__restrict__ terms,
                                                      int nt, double *__restrict__ part) {
    double v[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const cpx<R> a = psi[i];
        double acc = 0.0;
        for (int t = 0; t < nt; ++t) {
            const PauliTerm T = terms[t];
            const cpx<R> b = T.x ? psi[i ^ T.x] : a;
            // conj(b)·a
            const double re = (double)b.x * a.x + (double)b.y * a.y;
            const double im = (double)b.x * a.y - (double)b.y * a.x;
            acc += parity_sign(i & T.z) * (T.cre * re - T.cim * im);
        }
        v[0] += acc;
    }
    block_reduce_store<1>(v, part);
}

// Row i of H·ψ: Σ_t c_t i^{ny} (-1)^{pc((i^x)&z)} ψ[i^x], accumulated in FP64
// over the terms in table order, 4 terms at a time (their records and gathers
// in flight together).
template <typename R>
__device__ __forceinline__ void pauli_row(const cpx<R> *__restrict__ psi, uint64_t i,
                                          const PauliTerm *__restrict__ terms, int nt, double &ar, double &ai) {
    ar = 0.0;
    ai = 0.0;
    int t = 0;
    for (; t + 4 <= nt; t += 4) {
        PauliTerm T[4];
        cpx<R> b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) T[u] = terms[t + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) b[u] = psi[i ^ T[u].x];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double s = parity_sign((i ^ T[u].x) & T[u].z);
            ar += s * (T[u].cre * b[u].x - T[u].cim * b[u].y);
            ai += s * (T[u].cre * b[u].y + T[u].cim * b[u].x);
        }
    }
    for (; t < nt; ++t) {
        const PauliTerm T = terms[t];
        const uint64_t j = i ^ T.x;
        const cpx<R> b = psi[j];
        const double s = parity_sign(j & T.z);
        ar += s * (T.cre * b.x - T.cim * b.y);
        ai += s * (T.cre * b.y + T.cim * b.x);
    }
}

// out[i] = Σ_t c_t i^{ny} (-1)^{pc((i^x)&z)} ψ[i^x]
template <typename R>
__global__ void __launch_bounds__(256) k_pauli_apply(const cpx<R> *__restrict__ psi,
                                                     cpx<R> *__restrict__ out, uint64_t N,
                                                     const PauliTerm *__restrict__ terms, int nt) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        double ar, ai;
        pauli_row<R>(psi, i, terms, nt, ar, ai);
        out[i] = cpx<R>{(R)ar, (R)ai};
    }
}

// The same, and the block's partial of ⟨ψ|Hψ⟩ (re, im) from the FP64 rows
// before they are rounded to the state precision.
template <typename R>
__global__ void __launch_bounds__(256) k_pauli_apply_inner(const cpx<R> *__restrict__ psi,
                                                           cpx<R> *__restrict__ out, uint64_t N,
                                                           const PauliTerm *__restrict__ terms, int nt,
                                                           double *__restrict__ part) {
    double v[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        double ar, ai;
        pauli_row<R>(psi, i, terms, nt, ar, ai);
        out[i] = cpx<R>{(R)ar, (R)ai};
        const cpx<R> a = psi[i];
        v[0] += (double)a.x * ar + (double)a.y * ai;  // Re(conj(ψ_i) λ_i)
        v[1] += (double)a.x * ai - (double)a.y * ar;  // Im(conj(ψ_i) λ_i)
    }
    block_reduce_store<2>(v, part);
}

}  // namespace svh

using namespace svh;

static int check_state(const void *psi, int n, int dtype, int n_min = 1) {
    if (!psi) return fail(-1, "state pointer is NULL");
    if (n < n_min || n > kMaxQubits)
        return fail(-2, "n must be in [" + std::to_string(n_min) + ", 36], got " + std::to_string(n));
    if (dtype != SVH_C64 && dtype != SVH_C128) return fail(-3, "dtype must be SVH_C64 or SVH_C128");
    return 0;
}

static int upload_terms(int n, int n_terms, const uint64_t *xmask, const uint64_t *zmask,
                        const double *coeff, void *scratch, size_t scratch_bytes,
                        size_t reserve, PauliTerm **dev_terms, hipStream_t st) {
    if (n_terms < 0 || n_terms > 4096) return fail(-4, "n_terms must be in [0, 4096]");
    if (n_terms > 0 && (!xmask || !zmask || !coeff)) return fail(-1, "term arrays are NULL");
    const size_t need = reserve + sizeof(PauliTerm) * (size_t)n_terms;
    if (!scratch || scratch_bytes < need)
        return fail(-5, "scratch too small: need " + std::to_string(need) + " bytes");
    const uint64_t lim = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    static thread_local PauliTerm host[4096];
    for (int t = 0; t < n_terms; ++t) {
        if ((xmask[t] & ~lim) || (zmask[t] & ~lim)) return fail(-4, "Pauli mask beyond n bits");
        const int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
        // i^{ny}
        static const double pr[4] = {1, 0, -1, 0}, pi[4] = {0, 1, 0, -1};
        host[t] = PauliTerm{xmask[t], zmask[t], coeff[t] * pr[ny], coeff[t] * pi[ny]};
    }
    *dev_terms = reinterpret_cast<PauliTerm *>(static_cast<char *>(scratch) + reserve);
    if (n_terms == 0) return 0;
    // Synchronous with respect to the host buffer (pageable source).
    return hip_status(hipMemcpyAsync(*dev_terms, host, sizeof(PauliTerm) * n_terms,
                                     hipMemcpyHostToDevice, st),
                      "hipMemcpyAsync(terms)");
}

extern "C" {

const char *svh_last_error(void) { return g_last_error.c_str(); }

int svh_abi_version(void) { return SVH_ABI_VERSION; }

int svh_init_basis(void *psi, int n, int dtype, uint64_t idx, void *stream) {
    if (int e = check_state(psi, n, dtype)) return e;
    const uint64_t N = 1ull << n;
    if (idx >= N) return fail(-4, "basis index out of range");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(N, 256, 8192);
    if (dtype == SVH_C64)
        k_init_basis<float><<<g, 256, 0, st>>>((cpx<float> *)psi, N, idx);
    else
        k_init_basis<double><<<g, 256, 0, st>>>((cpx<double> *)psi, N, idx);
    return checked_launch("svh_init_basis");
}

int svh_apply_dense(void *psi, int n, int dtype, const int *bits, int k, const void *mat_dev,
                    void *stream) {
    if (int e = check_state(psi, n, dtype)) return e;
    if (k < 1 || k > 6 || k > n) return fail(-4, "gate size k must be in [1, min(6, n)]");
    if (!bits || !mat_dev) return fail(-1, "bits or matrix pointer is NULL");
    uint64_t mask = 0;
    int b[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < k; ++a) {
        if (bits[a] < 0 || bits[a] >= n) return fail(-4, "gate bit out of range");
        if ((mask >> bits[a]) & 1) return fail(-4, "repeated gate bit");
        mask |= 1ull << bits[a];
        b[a] = bits[a];
    }
    const uint64_t N = 1ull << n;
    hipStream_t st = (hipStream_t)stream;
    const unsigned grid = (unsigned)((N + 255) / 256);
    const int K = 1 << k;
    const int4 b03 = make_int4(b[0], b[1], b[2], b[3]);
    const int2 b45 = make_int2(b[4], b[5]);
    if (dtype == SVH_C64) {
        const size_t lds = sizeof(cpx<float>) * (K * K + 256);
        k_dense<float><<<grid, 256, lds, st>>>((cpx<float> *)psi, n, k, mask, b03, b45,
                                               (const cpx<float> *)mat_dev, N);
    } else {
        const size_t lds = sizeof(cpx<double>) * (K * K + 256);
        if (lds > 65536) {
            if (int e = ensure_lds_limit((const void *)k_dense<double>, (int)lds, "hipFuncSetAttribute(k_dense)"))
                return e;
        }
        k_dense<double><<<grid, 256, lds, st>>>((cpx<double> *)psi, n, k, mask, b03, b45,
                                                (const cpx<double> *)mat_dev, N);
    }
    return checked_launch("svh_apply_dense");
}

int svh_norm2(const void *psi, int n, int dtype, double *out_dev, void *scratch,
              size_t scratch_bytes, void *stream) {
    if (int e = check_state(psi, n, dtype, 0)) return e;
    if (!out_dev || !scratch) return fail(-1, "out or scratch is NULL");
    if (scratch_bytes < kReduceBlocks * sizeof(double)) return fail(-5, "scratch < 8 KiB");
    const uint64_t N = 1ull << n;
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(N, 256, kReduceBlocks);
    double *part = (double *)scratch;
    if (dtype == SVH_C64)
        k_norm2<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, N, part);
    else
        k_norm2<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, N, part);
    k_final_sum<1><<<1, 256, 0, st>>>(part, (int)g, out_dev);
    return checked_launch("svh_norm2");
}

int svh_scale_by_inv_norm(void *psi, int n, int dtype, const double *norm2_dev, void *stream) {
    if (int e = check_state(psi, n, dtype, 0)) return e;
    if (!norm2_dev) return fail(-1, "norm pointer is NULL");
    const uint64_t N = 1ull << n;
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(N, 256, 8192);
    if (dtype == SVH_C64)
        k_scale<float><<<g, 256, 0, st>>>((cpx<float> *)psi, N, norm2_dev);
    else
        k_scale<double><<<g, 256, 0, st>>>((cpx<double> *)psi, N, norm2_dev);
    return checked_launch("svh_scale_by_inv_norm");
}

int svh_inner(const void *a, const void *b, int n, int dtype, double *out_dev, void *scratch,
              size_t scratch_bytes, void *stream) {
    if (int e = check_state(a, n, dtype)) return e;
    if (!b || !out_dev || !scratch) return fail(-1, "NULL pointer");
    if (scratch_bytes < 2 * kReduceBlocks * sizeof(double)) return fail(-5, "scratch < 16 KiB");
    const uint64_t N = 1ull << n;
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(N, 256, kReduceBlocks);
    double *part = (double *)scratch;
    if (dtype == SVH_C64)
        k_inner<float><<<g, 256, 0, st>>>((const cpx<float> *)a, (const cpx<float> *)b, N, part);
    else
        k_inner<double><<<g, 256, 0, st>>>((const cpx<double> *)a, (const cpx<double> *)b, N, part);
    k_final_sum<2><<<1, 256, 0, st>>>(part, (int)g, out_dev);
    return checked_launch("svh_inner");
}

int svh_pauli_expect(const void *psi, int n, int dtype, int n_terms, const uint64_t *xmask,
                     const uint64_t *zmask, const double *coeff, double *out_dev, void *scratch,
                     size_t scratch_bytes, void *stream) {
    if (int e = check_state(psi, n, dtype)) return e;
    if (!out_dev) return fail(-1, "out pointer is NULL");
    hipStream_t st = (hipStream_t)stream;
    PauliTerm *terms = nullptr;
    const size_t reserve = kReduceBlocks * sizeof(double);
    if (int e = upload_terms(n, n_terms, xmask, zmask, coeff, scratch, scratch_bytes, reserve,
                             &terms, st))
        return e;
    const uint64_t N = 1ull << n;
    const unsigned g = grid_for(N, 256, kReduceBlocks);
    double *part = (double *)scratch;
    if (dtype == SVH_C64)
        k_pauli_expect<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, N, terms, n_terms, part);
    else
        k_pauli_expect<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, N, terms, n_terms, part);
    k_final_sum<1><<<1, 256, 0, st>>>(part, (int)g, out_dev);
    return checked_launch("svh_pauli_expect");
}

int svh_pauli_apply(const void *psi, void *out, int n, int dtype, int n_terms,
                    const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                    void *scratch, size_t scratch_bytes, void *stream) {
    if (int e = check_state(psi, n, dtype)) return e;
    if (!out) return fail(-1, "out pointer is NULL");
    if (out == psi) return fail(-4, "out must not alias psi");
    hipStream_t st = (hipStream_t)stream;
    PauliTerm *terms = nullptr;
    if (int e = upload_terms(n, n_terms, xmask, zmask, coeff, scratch, scratch_bytes, 0, &terms, st))
        return e;
    const uint64_t N = 1ull << n;
    const unsigned g = grid_for(N, 256, 16384);
    if (dtype == SVH_C64)
        k_pauli_apply<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, (cpx<float> *)out, N,
                                                terms, n_terms);
    else
        k_pauli_apply<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, (cpx<double> *)out, N,
                                                 terms, n_terms);
    return checked_launch("svh_pauli_apply");
}

int svh_pauli_apply_table(const void *psi, void *out, int n, int dtype, int n_terms, const void *terms_dev,
                          void *stream) {
    if (int e = check_state(psi, n, dtype)) return e;
    if (!out || (n_terms > 0 && !terms_dev)) return fail(-1, "NULL pointer");
    if (out == psi) return fail(-4, "out must not alias psi");
    if (n_terms < 0 || n_terms > 4096) return fail(-4, "n_terms must be in [0, 4096]");
    const uint64_t N = 1ull << n;
    const unsigned g = grid_for(N, 256, 16384);
    hipStream_t st = (hipStream_t)stream;
    const auto *terms = static_cast<const PauliTerm *>(terms_dev);
    if (dtype == SVH_C64)
        k_pauli_apply<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, (cpx<float> *)out, N, terms, n_terms);
    else
        k_pauli_apply<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, (cpx<double> *)out, N, terms, n_terms);
    return checked_launch("svh_pauli_apply_table");
}

int svh_pauli_apply_inner(const void *psi, void *out, int n, int dtype, int n_terms, const void *terms_dev,
                          double *out_dev, void *scratch, size_t scratch_bytes, void *stream) {
    if (int e = check_state(psi, n, dtype)) return e;
    if (!out || !out_dev || !scratch || (n_terms > 0 && !terms_dev)) return fail(-1, "NULL pointer");
    if (out == psi) return fail(-4, "out must not alias psi");
    if (n_terms < 0 || n_terms > 4096) return fail(-4, "n_terms must be in [0, 4096]");
    if (scratch_bytes < 2 * kReduceBlocks * sizeof(double)) return fail(-5, "scratch < 16 KiB");
    const uint64_t N = 1ull << n;
    const unsigned g = grid_for(N, 256, kReduceBlocks);
    hipStream_t st = (hipStream_t)stream;
    const auto *terms = static_cast<const PauliTerm *>(terms_dev);
    double *part = (double *)scratch;
    if (dtype == SVH_C64)
        k_pauli_apply_inner<float><<<g, 256, 0, st>>>((const cpx<float> *)psi, (cpx<float> *)out, N, terms,
                                                      n_terms, part);
    else
        k_pauli_apply_inner<double><<<g, 256, 0, st>>>((const cpx<double> *)psi, (cpx<double> *)out, N, terms,
                                                       n_terms, part);
    k_final_sum<2><<<1, 256, 0, st>>>(part, (int)g, out_dev);
    return checked_launch("svh_pauli_apply_inner");
}

}  // extern "C"
