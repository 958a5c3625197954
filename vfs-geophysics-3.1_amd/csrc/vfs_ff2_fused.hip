// Fused Formfunction_2 (momentum.c:456-1979): placeholder that routes to the
// reference-structured passes until the k-marching kernel lands.
#include "vfs_internal.h"

int vfs_ff2_passes(vfs_ctx *c, const double *ucont, const double *ucat, const double *nu_t, double *rhs, double scale);

int vfs_ff2_fused(vfs_ctx *c, const double *ucont, const double *ucat, const double *nu_t, double *rhs, double scale)
{
    return vfs_ff2_passes(c, ucont, ucat, nu_t, rhs, scale);
}
