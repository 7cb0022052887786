/* Types and constants of the qaravan HIP statevector library's C ABI: the
 * gate-program layout, pass/stage/op records and batch item tables.  The
 * entry points are declared in qaravan_hip.h, which includes this file.
 * Code generated for specialised passes depends on this file only (not on
 * the entry-point declarations), so its content keys the kernel cache. */
#ifndef QARAVAN_HIP_TYPES_H
#define QARAVAN_HIP_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVH_ABI_VERSION 1

enum { SVH_C64 = 0, SVH_C128 = 1 };

/* ---- gate program ---------------------------------------------------------
 * A program is one contiguous byte buffer: svh_prog_header_t, then the pass,
 * stage, op, grad-slot (int32) and matrix-pool (double) arrays at the byte
 * offsets the header gives.  The caller builds it on the host, copies it to
 * device memory, and passes both copies to svh_run_program / svh_run_adjoint:
 * the host copy drives the launches, the device copy is read by the kernels.
 */
#define SVH_PROG_MAGIC 0x53564850u /* 'SVHP' */
#define SVH_MAX_TILE_BITS 16
#define SVH_MAX_REG_BITS 5

enum {
    SVH_PASS_TILE = 1,      /* fused pass over tiles of 2^L amplitudes            */
    SVH_PASS_DENSE = 2,     /* one dense k-qubit gate (k <= 6) from the pool        */
    SVH_PASS_NORMALIZE = 3  /* ψ /= ‖ψ‖₂ (FP64 norm), used after Reset projectors   */
};

enum {
    SVH_OP_MAT1 = 1, /* dense 2x2 on register bit a                               */
    SVH_OP_MAT2 = 2, /* dense 4x4 on register bits (a, b), a > b, a = row MSB       */
    SVH_OP_CX = 3,   /* controlled-X: a = target bit, b = control bit              */
    SVH_OP_CZ = 4,   /* diag(1,1,1,-1) on (a, b)                                   */
    SVH_OP_ROT1 = 5, /* exp(-i·s·θ[param]·P), P = pauli (1 X, 2 Y, 3 Z), on a       */
    SVH_OP_ROT2 = 6  /* exp(-i·s·θ[param]·P⊗P) on (a, b)                           */
};

enum {
    SVH_OPF_NEG = 1, /* s = -1 (the adjoint sweep applies U†)                     */
    SVH_OPF_GRAD = 2 /* before applying: grad_slot += 2·Im⟨λ|P|ψ⟩ (adjoint only)   */
};

typedef struct {
    uint32_t magic;
    int32_t version;
    int32_t n_qubits;
    int32_t n_passes;
    int32_t n_stages;
    int32_t n_ops;
    int32_t n_grad_slots;
    int32_t n_pool; /* doubles in the matrix pool */
    int64_t off_passes, off_stages, off_ops, off_grad, off_pool;
    int64_t n_params; /* length of θ read by ROT ops (0 if none)                  */
    /* 0, or the offset of int32 perm[n_qubits]: the program's passes move
     * logical address bit q to physical bit perm[q]; the run ends with one
     * permutation back to the natural order (needs scratch for one state). */
    int64_t off_perm;
    /* 0, or (forward programs with off_perm set) the offset of int32 count
     * (1..8) followed by count (control, target) pairs of physical address
     * bits: CNOTs that run, in this order, after the last pass and before the
     * final permutation, which applies them while it gathers (a trailing pass
     * of CNOTs only costs no sweep of its own). */
    int64_t off_cx;
} svh_prog_header_t; /* 96 bytes */

typedef struct {
    int32_t kind;        /* SVH_PASS_*                                              */
    int32_t L;           /* tile bits (TILE) or gate qubits k (DENSE)              */
    int32_t tb;          /* log2 threads per workgroup (TILE)                       */
    int32_t rb;          /* register bits per thread (TILE); L = tb + rb            */
    int32_t stage_begin; /* first stage (TILE)                                       */
    int32_t n_stages;
    int32_t grad_begin;  /* first grad slot of this pass (adjoint)                  */
    int32_t n_grad;
    int64_t pool_off;    /* DENSE: offset (doubles) of the 2^k x 2^k complex matrix  */
    int32_t flags;       /* SVH_PASSF_*                                              */
    int32_t pad;
    int32_t bits[SVH_MAX_TILE_BITS]; /* TILE: global bit of tile-local bit i (i < L);
                                        DENSE: global bit of matrix qubit i (i = 0 is
                                        the most significant row bit)                */
} svh_pass_t; /* 112 bytes */

enum {
    /* TILE (forward only): persistent workgroups; the next tile is copied
     * HBM -> LDS asynchronously while the current one is computed.  Needs
     * tile-local bit 0 = address bit 0. */
    SVH_PASSF_PREFETCH = 1,
    /* TILE, per-pass compiled forward kernels: every LDS re-distribution moves
     * the tile in two halves (HALF) or four quarters (QUARTER), split on
     * register bits both layouts share, so the pass needs LDS for 2^(L-1) or
     * 2^(L-2) amplitudes instead of 2^L. */
    SVH_PASSF_HALF_LDS = 2,
    SVH_PASSF_QUARTER_LDS = 4,
    /* TILE, per-pass compiled forward kernels: the tile is stored through the
     * stage record that follows the pass's n_stages records; its global bits
     * permute the tile's bits (a qubit remap within the tile). */
    SVH_PASSF_REMAP_STORE = 8,
    /* TILE: run by the generic kernels (one kernel for every pass structure:
     * stage records walked at run time, each stage dispatched to the compiled
     * pattern library or interpreted); 3 or 4 register bits, tb >= 6.  In an
     * adjoint sweep every wave writes its own row of gradient partials, so a
     * pass has 2^(n-L) * 2^(tb-6) rows of n_grad doubles. */
    SVH_PASSF_GENERIC = 16,
    /* TILE, first pass of a forward program, per-pass compiled kernel only:
     * the pass does not read the state; it starts from |0…0⟩, generated in
     * its registers (the caller skips the initialisation sweep). */
    SVH_PASSF_GEN_ZERO = 32
};

/* A stage's layout: thread t holds registers j = 0..2^rb-1 at tile-local index
 * insert-zeros(t at rsorted) | loc(j), i.e. at global amplitude
 * base | Σ_i bit_i(t)·2^tgbit[i] | goff[j].  The tables are precomputed by the
 * host so the kernel does no bit arithmetic per amplitude.  The op-interpreting
 * kernels run rb = 4; the per-pass compiled kernels (svh_sp_*) run rb = 4 or 5. */
typedef struct {
    int64_t goff[32];    /* global amplitude offset of register j                 */
    int32_t lsw[32];     /* swizzled tile-local (LDS) offset of register j         */
    int32_t rbits[8];    /* tile-local bit held in register-index bit k (k < rb)   */
    int32_t rsorted[8];  /* rbits[0..rb) in ascending order                         */
    int32_t tgbit[12];   /* global bit of thread-index bit i (i < tb)              */
    int32_t op_begin, op_end;
    int32_t pattern;     /* generic kernels: id of this stage's op sequence in the
                            compiled pattern library of the pass's (rb, direction),
                            or -1: the stage is interpreted op by op            */
    int32_t pad;
} svh_stage_t; /* 512 bytes */

typedef struct {
    int32_t kind;        /* SVH_OP_*                                                */
    int32_t a, b;        /* register-index bits                                      */
    int32_t param;       /* θ index (ROT*)                                           */
    int32_t pauli;       /* 1 X, 2 Y, 3 Z (ROT*)                                     */
    int32_t flags;       /* SVH_OPF_*                                                */
    int32_t grad_slot;   /* pass-local gradient slot (SVH_OPF_GRAD); consecutive within a stage */
    int32_t pad;
    double m[32];        /* complex row-major matrix (re, im pairs)                  */
    float mf[32];        /* the same matrix rounded to float (used by SVH_C64)       */
} svh_op_t; /* 416 bytes */

/* ---- batches of circuits ----------------------------------------------------
 * Many independent circuits (different n, depth, θ) advance together: round r
 * of a batch runs pass r of every circuit that has one, in ONE launch per
 * (tile shape, mode), so small circuits fill the GPU together.  The caller
 * (host planner) builds the item tables once, in device memory; a whole batch
 * (init, trig, forward rounds, H·ψ and energies, adjoint rounds with gradient
 * reduction) is a fixed launch sequence that can be captured in a HIP graph. */
typedef struct {
    uint64_t psi, lam;          /* device state / λ pointers (lam = 0 forward)      */
    uint64_t pass, stages, ops; /* device program pointers: this circuit's pass r   */
    uint64_t trig;              /* device (cos, sin) table of the circuit          */
    uint64_t gpart;             /* device per-workgroup gradient partials           */
    int64_t wg_begin;           /* first workgroup of the item in the launch         */
    int32_t n, pad;
} svh_batch_item_t; /* 72 bytes */

typedef struct {
    uint64_t part;              /* device partial sums: row b of ng values          */
    uint64_t slot_param;        /* device int32 slot -> output index                 */
    uint64_t out;               /* device float64 outputs (grad or energy)           */
    uint64_t tstamp;            /* device uint64 completion timestamp or 0           */
    int32_t nblocks, ng;        /* partial rows, slots                                */
    int64_t slot_begin;         /* first launch block of the item                     */
} svh_batch_red_t; /* 48 bytes */

typedef struct {
    uint64_t psi, lam;          /* device ψ (read) and λ = Hψ (written)             */
    uint64_t terms;             /* device term table (32 B per term: x, z, re, im)    */
    uint64_t epart;             /* device FP64 per-block partials of Re⟨ψ|λ⟩         */
    int64_t block_begin;        /* first launch block of the item                     */
    int32_t n, nterms;
} svh_batch_ham_t; /* 48 bytes */

/* ---- persistent batch ---------------------------------------------------------
 * One launch runs a whole batch (forward passes, λ = Hψ with the energy,
 * adjoint passes, gradient sums) as a list of tasks.  Each task covers
 * `count` units of one descriptor: tiles of a pass, 1024-amplitude blocks of
 * the Hamiltonian step, or slots of a reduction.  A descriptor waits until
 * counter `dep` reaches `dep_need` (the units of the step it reads) and adds
 * its finished units to counter `done`, so every circuit advances as soon as
 * its own previous step is complete.  Tasks sit in per-queue lists in an
 * order where every dependency comes earlier in the same queue. */
enum {
    SVH_PT_FWD = 1, /* tile units of a forward pass                               */
    SVH_PT_ADJ = 2, /* tile units of an adjoint pass (ψ and λ, gradient rows)      */
    SVH_PT_HAM = 3, /* blocks of λ = Σ c_t P_t ψ plus per-block Re⟨ψ|λ⟩ partials   */
    SVH_PT_RED = 4, /* slot sums of partial rows, fixed order                      */
    SVH_PT_SCT = 5  /* one unit: out[slot_param[s]] += slotsum[s] for s ascending   */
};

typedef struct {
    int32_t kind;       /* SVH_PT_*                                                  */
    int32_t n;          /* qubits of the circuit                                       */
    int32_t dep;        /* counter to wait for, or -1                                  */
    int32_t dep_need;   /* value that counter must reach                               */
    int32_t done;       /* counter that receives the finished units, or -1             */
    int32_t units;      /* tiles / blocks / slots                                      */
    int32_t ng;         /* ADJ, RED: gradient slots (partial-row length); HAM: terms  */
    int32_t rows;       /* RED: partial rows                                           */
    int32_t fin;        /* RED, SCT: counter of a circuit's finished sums, or -1       */
    int32_t fin_need;   /* the value of `fin` once the circuit's gradient is complete  */
    uint64_t psi, lam;  /* FWD, ADJ, HAM: state and adjoint vectors                    */
    uint64_t pass;      /* FWD, ADJ: the pass record                                   */
    uint64_t stages;    /* FWD, ADJ: the pass's first stage record                     */
    uint64_t ops;       /* FWD, ADJ: the program's op array                            */
    uint64_t trig;      /* FWD, ADJ: (cos, sin) table of the circuit                   */
    uint64_t part;      /* ADJ, RED: partial rows; HAM: per-block energy partials      */
    uint64_t out;       /* RED, SCT: float64 outputs; HAM: the term table              */
    uint64_t slot_param;/* RED, SCT: int32 output index per slot                       */
    uint64_t slotsum;   /* RED: 0 = store the sum at out[slot_param[s]]; else slotsum[s]
                           (SCT then adds them in slot order)                           */
    uint64_t tstamp;    /* RED, SCT: uint64 completion time of the circuit, or 0       */
} svh_pdesc_t; /* 128 bytes */

typedef struct {
    int32_t desc, first, count, pad;
} svh_ptask_t; /* 16 bytes */

#ifdef __cplusplus
}
#endif

#endif /* QARAVAN_HIP_TYPES_H */
