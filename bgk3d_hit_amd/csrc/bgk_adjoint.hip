// bgk_adjoint.hip -- discrete adjoint of the fused stream-collide.
#include "bgk_internal.h"

namespace bgk {

int forward_ke(bgk_ctx *c, const double *u0_dev, int nsteps, double *J) {
    (void)c; (void)u0_dev; (void)nsteps; (void)J;
    return fail(BGK_ESTATE, "forward_ke: not available in this build");
}

int adjoint_ke(bgk_ctx *c, const double *u0_dev, int nsteps, int ckpt_every, int levels,
               double *J, double *grad_dev, double stats[4]) {
    (void)c; (void)u0_dev; (void)nsteps; (void)ckpt_every; (void)levels; (void)J;
    (void)grad_dev; (void)stats;
    return fail(BGK_ESTATE, "adjoint_ke: not available in this build");
}

}  // namespace bgk
