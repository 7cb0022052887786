/*
 * qaravan_hip.h — C ABI of the MI355X (gfx950) statevector backend.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns
 * an int status: 0 = success, < 0 = argument error (message in svh_last_error()),
 * > 0 = the hipError_t that a HIP call returned.  All device buffers (state
 * vectors, programs, scratch) are allocated by the caller; the library keeps no
 * device memory of its own.  Work is enqueued on the given stream (NULL = the
 * default stream) and entry points return without synchronising unless stated.
 *
 * Qubit/bit convention: the state is a flat vector of 2^n complex amplitudes,
 * interleaved (re, im) in float (dtype SVH_C64) or double (SVH_C128).  The
 * library works in ADDRESS BITS: amplitude index i has bit b = (i >> b) & 1.
 * qaravan's big-endian qubit q is address bit n-1-q; the host side converts.
 *
 * Reference interfaces these replace (qaravan 0.1.57, src/qaravan/...):
 *   svh_init_basis        tensorQ/statevector_sim.py:31-48  StatevectorSim.initialize_state
 *   svh_apply_dense       tensorQ/statevector_sim.py:50-52,124-140  apply_gate -> op_action
 *   svh_run_program       core/base_sim.py:30-53  BaseSim.run (the per-layer/per-gate loop)
 *   svh_sp_*              the same loop, with each fused pass compiled for its gate list
 *   svh_norm2             tensorQ/statevector_sim.py:107-109  normalize_state (norm)
 *   svh_scale             tensorQ/statevector_sim.py:107-109  normalize_state (divide)
 *   svh_pauli_expect      tensorQ/statevector_sim.py:217-234  pauli_expectation_sv, summed
 *                         over Hamiltonian terms (core/hamiltonians.py:19-88)
 *   svh_pauli_apply       .dev/test_hamiltonians.ipynb:709-712 ham_action (H·ψ via op_action)
 *   svh_pauli_apply_inner the same H·ψ with ⟨ψ|H·ψ⟩ (the energy a gradient step needs) in one sweep
 *   svh_inner             tensorQ/statevector_sim.py:155-180 partial_overlap with no kept
 *                         sites, i.e. ⟨a|b⟩ (applications/compilation.py:86 torch.dot)
 *   svh_run_adjoint       applications/compilation.py:73-133 (cost.backward() through the
 *                         torch backend); same gradient, computed by the adjoint method
 *   svh_apply_block5      no counterpart: op_action applies a 5-qubit gate with one
 *                         tensordot; here it is an FP32 MFMA GEMM (c64)
 *   svh_apply_block5x2    two such op_action calls (statevector_sim.py:50-52) on disjoint
 *                         qubits of one layer, fused into one sweep over the state
 *   svh_permute_bits      no counterpart: the reference keeps qubit q at axis q; remapped
 *                         programs move qubits between address bits and end with this
 *   svh_batch_*, svh_spb_* tensorQ/monte_carlo_sim.py:14-35 (many small StatevectorSim runs,
 *                         one after another); here a whole batch of parametrised circuits
 *                         with their adjoint gradients runs as one launch sequence
 *   svh_batch_persistent  the same many runs (and applications/compilation.py:73-133 once
 *                         per circuit), as one persistent launch with a device task queue
 *   svh_timestamp         no counterpart (device clock for per-circuit latencies)
 *   svh_marginal_probs    tensorQ/statevector_sim.py:185-187 measure_sv (diag of rdm_from_sv)
 *   svh_sample_cdf        tensorQ/statevector_sim.py:188-193 np.random.choice(outcomes, p=probs)
 *   svh_collapse_gather   tensorQ/statevector_sim.py:195-215 measure_and_collapse_sv (projection)
 *   svh_partial_overlap   tensorQ/statevector_sim.py:155-183 partial_overlap / rdm_from_sv,
 *                         and the environment tensors of tensorQ/environments.py:7-23
 */
#ifndef QARAVAN_HIP_H
#define QARAVAN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#include "qaravan_hip_types.h"

/* ---- status ---------------------------------------------------------------- */
const char *svh_last_error(void);
int svh_abi_version(void);

/* ψ = |idx⟩ (all zeros, amplitude idx = 1). */
int svh_init_basis(void *psi, int n, int dtype, uint64_t idx, void *stream);

/* Apply a dense k-qubit gate (1 <= k <= 6) in place.  bits[a] = address bit of
 * matrix qubit a (a = 0 most significant); mat_dev = device 2^k x 2^k complex
 * matrix, row-major, in the state's precision. */
int svh_apply_dense(void *psi, int n, int dtype, const int *bits, int k,
                    const void *mat_dev, void *stream);

/* Apply a dense 5-qubit gate with the FP32 matrix cores (SVH_C64 only, n >= 12).
 * bits[j] = address bit of matrix row bit j (j = 0 least significant);
 * u_dev = device 32x32 complex64 matrix, row-major.  Programs route their
 * 5-qubit dense passes here for complex64 states. */
int svh_apply_block5(void *psi, int n, int dtype, const int *bits, const void *u_dev, void *stream);

/* Apply two dense 5-qubit gates on disjoint bits (they commute) in one sweep,
 * both on the FP32 matrix cores (SVH_C64 only, n >= 12).  bits_a / bits_b and
 * ua_dev / ub_dev as for svh_apply_block5.  Programs route two consecutive
 * complex64 5-qubit dense passes on disjoint bits here. */
int svh_apply_block5x2(void *psi, int n, int dtype, const int *bits_a, const void *ua_dev,
                       const int *bits_b, const void *ub_dev, void *stream);

/* Out-of-place bit permutation: dst[i'] = src[i] where bit b of i becomes bit
 * perm[b] of i' (perm is a host array of n distinct bits; 8 <= n <= 36).  A
 * program compiled with qubit remapping ends with this step inside
 * svh_run_program_sp, through scratch, so its result is in natural order. */
int svh_permute_bits(const void *src, void *dst, int n, int dtype, const int *perm, void *stream);

/* Bytes of device scratch svh_run_program / svh_run_adjoint need for `prog`. */
int svh_program_scratch_bytes(const void *prog_host, int dtype, size_t *bytes);

/* Run every pass of a program on ψ in place.  theta_dev: device float64 angles
 * for ROT ops (may be NULL if the program has none); cos/sin of every angle are
 * evaluated once per call in FP64 and rounded to the state's precision. */
int svh_run_program(void *psi, int n, int dtype, const void *prog_host,
                    const void *prog_dev, const double *theta_dev, void *scratch,
                    size_t scratch_bytes, void *stream);

/* Allocation-free svh_run_program: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops).  The scratch buffer is a fixed 16 MiB device array inside
 * the library (each GPU it runs on holds its own copy until the process ends),
 * so the caller passes only ψ and the program.  Programs whose scratch need
 * (svh_program_scratch_bytes) exceeds 16 MiB are refused with -5; programs
 * with compiled-only passes are refused as by svh_run_program.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU race. */
int svh_run_program_static(void *psi, int n, int dtype, const void *prog_host,
                           const void *prog_dev, const double *theta_dev, void *stream);

/* Specialised programs: code objects (hipcc --genco) holding kernels named
 * svh_sp_<p>, one per tile pass p they cover, generated for one program with
 * every op spelled out (qaravan_amd/specialize.py).  svh_sp_load resolves the
 * kernels of n_passes passes across the given code-object files;
 * svh_run_program_sp runs the program launching them where present and the
 * generic kernels elsewhere.  The handle owns the loaded modules.  Kernel
 * signature: svh_sp_<p>(psi, n, pass, stages, ops, trig); in a program with a
 * final permutation (header off_perm), the last pass's kernel takes an output
 * pointer after psi and stores there (the scratch tail the permutation reads). */
int svh_sp_load(const char *const *paths, int npaths, int n_passes, void **handle);
int svh_sp_free(void *handle);
int svh_run_program_sp(void *handle, void *psi, int n, int dtype, const void *prog_host,
                       const void *prog_dev, const double *theta_dev, void *scratch,
                       size_t scratch_bytes, void *stream);
/* Adjoint sweep with the kernels svh_spd_<p> of a specialised adjoint program. */
int svh_run_adjoint_sp(void *handle, void *psi, void *lam, int n, int dtype, const void *prog_host,
                       const void *prog_dev, const double *theta_dev, double *grad_dev, void *scratch,
                       size_t scratch_bytes, void *stream);

/* Adjoint sweep: the program holds the reversed, daggered gate list with
 * SVH_OPF_GRAD on rotations.  Walks ψ and λ back together and ADDS
 * 2·Im⟨λ_k|P_k|ψ_k⟩ into grad_dev[param] (device float64, caller zeroes it). */
int svh_run_adjoint(void *psi, void *lam, int n, int dtype, const void *prog_host,
                    const void *prog_dev, const double *theta_dev, double *grad_dev,
                    void *scratch, size_t scratch_bytes, void *stream);

/* out_dev[0] = Σ|ψ_i|² accumulated in FP64.  scratch >= 8·1024 bytes. */
int svh_norm2(const void *psi, int n, int dtype, double *out_dev, void *scratch,
              size_t scratch_bytes, void *stream);

/* ψ *= 1/sqrt(norm2_dev[0]) (norm2_dev: device double written by svh_norm2). */
int svh_scale_by_inv_norm(void *psi, int n, int dtype, const double *norm2_dev,
                          void *stream);

/* Pauli sums: term t is (xmask[t], zmask[t], coeff[t]) with P_t|i⟩ =
 * i^{ny_t}·(-1)^{popcount(i & zmask)}·|i ^ xmask⟩, ny_t = popcount(x & z).
 * Term arrays are HOST arrays (n_terms <= 4096), coeff real.
 * svh_pauli_expect: out_dev[0] = Σ_t c_t Re⟨ψ|P_t|ψ⟩ (FP64 accumulation). */
int svh_pauli_expect(const void *psi, int n, int dtype, int n_terms,
                     const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                     double *out_dev, void *scratch, size_t scratch_bytes, void *stream);

/* out = Σ_t c_t P_t ψ  (out must not alias psi); the term table is staged in
 * scratch (>= 32·n_terms bytes). */
int svh_pauli_apply(const void *psi, void *out, int n, int dtype, int n_terms,
                    const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                    void *scratch, size_t scratch_bytes, void *stream);

/* One batch round: k_tile over every item (items_dev sorted by wg_begin,
 * total_wg workgroups), all with the same tb and mode (dual = adjoint). */
int svh_batch_tile(const void *items_dev, int nitems, int64_t total_wg, int dtype, int tb, int dual,
                   int max_ngrad, void *stream);

/* Sum partial rows per (item, slot) in fixed order and add into
 * out[slot_param[slot]]; writes a completion timestamp (s_memrealtime, 100 MHz)
 * with atomicMax when tstamp != 0.  total_blocks = Σ ng. */
int svh_batch_reduce(const void *red_dev, int nitems, int64_t total_blocks, void *stream);

/* λ = Hψ for every item and per-block partials of Re⟨ψ|λ⟩ (1024 amplitudes
 * per block; total_blocks = Σ 2^n / 1024). */
int svh_batch_hamiltonian(const void *ham_dev, int nitems, int64_t total_blocks, int dtype, void *stream);

/* buf[i] = (i in offsets) ? 1 : 0 over a flat buffer holding many states
 * (offsets sorted ascending, device int64). */
int svh_batch_init(void *buf, int64_t total, const int64_t *offsets_dev, int noffsets, int dtype,
                   void *stream);

/* (cos θ, sin θ) for a concatenated θ vector, rounded to the state precision. */
int svh_batch_trig(const double *theta_dev, int64_t np, void *trig_dev, int dtype, void *stream);

/* Batch kernels compiled per pass structure (svh_spb_*): load the named
 * kernels from code objects (handle freed with svh_sp_free), then launch
 * kernel ``index`` over an item table as svh_batch_tile does: workgroup b
 * runs item i with wg_begin[i] <= b < wg_begin[i+1], every item's pass having
 * the structure the kernel was compiled for. */
int svh_spb_load(const char *const *paths, int npaths, const char *const *names, int nnames, void **handle);
int svh_spb_launch(void *handle, int index, const void *items_dev, int nitems, int64_t total_wg, int tb,
                   size_t lds_bytes, void *stream);

/* Persistent batch: ONE launch (after three small fill kernels that zero the
 * counters, queue heads and error flag) runs every task of a batch.  descs_dev:
 * svh_pdesc_t array; tasks_dev: svh_ptask_t array holding nq queues back to
 * back, queue q = tasks [qbeg[q], qbeg[q+1]) (qbeg_dev: nq + 1 int32, device);
 * heads_dev: nq int32 scratch; counters_dev: ncounters int32 scratch named by
 * the descriptors' dep/done/fin fields; err_dev: one int32, set to 1 when a
 * dependency wait exceeded its bound, 2 when a tile pass does not fit the
 * workgroup (more than `threads` threads, register bits other than 3, or more
 * amplitudes than the λ offset; its units are skipped; the results are then
 * invalid).  Every
 * descriptor's dependencies must come earlier in its own queue.  lds_bytes:
 * the LDS a task needs; the launch takes at least the ψ and λ tiles of the
 * largest adjoint tile (λ at a fixed offset of threads·8 complex64 or
 * threads·4 complex128 amplitudes; tiles with more amplitudes are refused);
 * threads: workgroup size, 512 (0 means 512) or, for complex64, 256; a tile
 * pass's 2^tb must not exceed it.
 * max_workgroups: cap on the resident grid (0 = fill the device).
 * xcd_local != 0 asks for XCD-local synchronisation: a workgroup serves only
 * the queue of the XCD it runs on and dependencies are made visible through
 * that XCD's shared L2 (no device-wide write-back / invalidation per task);
 * it is used only when the full grid covers exactly XCDs 0..nq-1 (checked
 * once per device and grid shape by a small census kernel with a temporary
 * allocation, outside graph capture), otherwise workgroups help other queues
 * with device-scope release/acquire.  *used_mode (optional host int) = 1 for
 * XCD-local, 0 otherwise.  stats_dev (optional, 5 uint64 on the device, which
 * the caller zeroes): every workgroup adds its dependency-wait time, its task
 * time (claim to counter update, waits included), in 100 MHz ticks, its task
 * count, the number of its waits that had to spin, and 1 (the grid size).  Tile passes must be
 * generic passes with 3 register bits.  Calls that overlap on one set of
 * counters race. */
int svh_batch_persistent(const void *descs_dev, const void *tasks_dev, const int32_t *qbeg_dev, int nq,
                         int32_t *heads_dev, int32_t *counters_dev, int ncounters, int32_t *err_dev, int dtype,
                         size_t lds_bytes, int threads, int max_workgroups, int xcd_local, int *used_mode,
                         uint64_t *stats_dev, void *stream);

/* out_dev[0] = wall clock (s_memrealtime, 100 MHz ticks), for batch latencies. */
int svh_timestamp(uint64_t *out_dev, void *stream);

/* The same with the term table already in device memory (32 bytes per term:
 * uint64 xmask, uint64 zmask, double Re(c·i^{ny}), double Im(c·i^{ny})).  No host
 * copies, so it can be recorded in a HIP graph. */
int svh_pauli_apply_table(const void *psi, void *out, int n, int dtype, int n_terms, const void *terms_dev,
                          void *stream);

/* svh_pauli_apply_table and ⟨ψ|Hψ⟩ in one sweep: out = H·ψ as above, and
 * out_dev[0..1] = (re, im) of Σ_i conj(ψ_i)·(Hψ)_i, accumulated in FP64 from
 * the rows before they are rounded to the state precision (the loss value of
 * applications/compilation.py-style energies, tensorQ/statevector_sim.py:217-234
 * summed over the Hamiltonian's terms).  scratch >= 16 KiB. */
int svh_pauli_apply_inner(const void *psi, void *out, int n, int dtype, int n_terms, const void *terms_dev,
                          double *out_dev, void *scratch, size_t scratch_bytes, void *stream);

/* out_dev[0..1] = ⟨a|b⟩ (re, im) accumulated in FP64. */
int svh_inner(const void *a, const void *b, int n, int dtype, double *out_dev,
              void *scratch, size_t scratch_bytes, void *stream);

/* ---------------------------------------------------------------- measurement
 * A set of k "kept" qubits is passed as k distinct address bits, most
 * significant first: kept[0] carries bit k-1 of the kept index m.  The other
 * n-k address bits carry the rest index r in ascending order (big-endian over
 * the remaining qubits).  Sums accumulate in FP64 in a fixed order. */
enum { SVH_MEAS_PROBS = 1, SVH_MEAS_SAMPLE = 2, SVH_MEAS_OVERLAP = 3 };

/* Scratch bytes that svh_marginal_probs / svh_sample_cdf / svh_partial_overlap
 * (what = SVH_MEAS_*) need for n qubits and k kept ones (0 = invalid). */
size_t svh_measure_scratch_bytes(int what, int n, int k);

/* probs[m] = Σ_r |ψ(m, r)|² (FP64, 2^k entries). */
int svh_marginal_probs(const void *psi, int n, int dtype, const int *kept, int k, double *probs,
                       void *scratch, size_t scratch_bytes, void *stream);

/* Inverse-CDF sampling, the algorithm of numpy's np.random.choice(a, p=probs):
 * cdf = cumsum(probs), idx[s] = first i with cdf[i]/cdf[M-1] > uniforms[s]
 * (searchsorted, side='right').  total[0] = cdf[M-1] (for the caller's sum
 * check).  M = 2^k; the uniforms come from the caller. */
int svh_sample_cdf(const double *probs, int k, const double *uniforms, int64_t shots, int64_t *idx,
                   double *total, void *scratch, size_t scratch_bytes, void *stream);

/* out[r] = ψ(outcome, r): the 2^(n-k) amplitudes whose kept bits equal outcome
 * (same dtype as ψ; normalise with svh_norm2 + svh_scale_by_inv_norm). */
int svh_collapse_gather(const void *psi, int n, int dtype, const int *kept, int k, uint64_t outcome, void *out,
                        void *stream);

/* out[i*2^k + j] = Σ_r a(i, r)·conj(b(j, r)) as complex128, k ≤ 6. */
int svh_partial_overlap(const void *a, const void *b, int n, int dtype, const int *kept, int k, void *out,
                        void *scratch, size_t scratch_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QARAVAN_HIP_H */
