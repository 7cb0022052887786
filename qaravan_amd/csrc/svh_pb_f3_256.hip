// Persistent batch kernel, float states, 256-thread workgroups (four resident
// per CU), 3 register bits (see svh_gtile_kern.h).
#include "svh_gtile_kern.h"

namespace svh {

const void *persist_kernel_f3_256() { return (const void *)k_batch_persist<float, 256, 3>; }

int persist_launch_f3_256(unsigned grid, unsigned threads, size_t lds, hipStream_t st, const svh_pdesc_t *D,
                          const svh_ptask_t *T, const int32_t *qbeg, int nq, int32_t *heads, int32_t *ctr,
                          int32_t *err, int mode, unsigned long long *stats) {
    if (threads != 256) return fail(-8, "persistent batch: workgroup size does not match the kernel");
    k_batch_persist<float, 256, 3><<<grid, 256, lds, st>>>(D, T, qbeg, nq, heads, ctr, err, mode, stats);
    return checked_launch("k_batch_persist");
}

}  // namespace svh
