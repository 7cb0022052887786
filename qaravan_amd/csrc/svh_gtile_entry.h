// Entry points of the per-instantiation files of the generic tile kernels
// (svh_gt_<dtype><rb>.hip: one pass; svh_pb_<dtype><rb>.hip: the persistent
// batch), called by the dispatchers in svh_gtile.hip.
#pragma once
#include "svh_common.h"

namespace svh {

#define SVH_GT_DECL(X)                                                                                          \
    int gtile_launch_##X(const void *psi, void *out, void *lam, int n, const svh_pass_t &hp, const svh_pass_t *dp, \
                         const svh_stage_t *dst, const svh_op_t *dop, const void *trig, double *gpart,             \
                         hipStream_t stream);
SVH_GT_DECL(f3)
SVH_GT_DECL(f4)
SVH_GT_DECL(d3)
SVH_GT_DECL(d4)
#undef SVH_GT_DECL

#define SVH_PB_DECL(X)                                                                                          \
    const void *persist_kernel_##X();                                                                           \
    int persist_launch_##X(unsigned grid, unsigned threads, size_t lds, hipStream_t st, const svh_pdesc_t *D,    \
                           const svh_ptask_t *T, const int32_t *qbeg, int nq, int32_t *heads, int32_t *ctr,    \
                           int32_t *err, int mode, unsigned long long *stats);
SVH_PB_DECL(f3)
SVH_PB_DECL(d3)
SVH_PB_DECL(f3_256)
#undef SVH_PB_DECL

}  // namespace svh
