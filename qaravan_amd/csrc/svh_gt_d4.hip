// Generic tile pass, double states, 4 register bits (see svh_gtile_kern.h).
#include "svh_gtile_kern.h"

namespace svh {

int gtile_launch_d4(const void *psi, void *out, void *lam, int n, const svh_pass_t &hp, const svh_pass_t *dp,
                     const svh_stage_t *dst, const svh_op_t *dop, const void *trig, double *gpart,
                     hipStream_t stream) {
    return lam ? launch_gtile<double, 4, true>(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream)
               : launch_gtile<double, 4, false>(psi, out, lam, n, hp, dp, dst, dop, trig, gpart, stream);
}

}  // namespace svh
