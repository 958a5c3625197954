// RungeKutta (rhs.c:1578-1700): the explicit momentum step of Flow_Solver in
// the reference's order, with the boundary work around it.
//
// Before the stages: InflowFlux on an inlet at k = 0, FormBCS's outlet
// branch with outflow_scale = 0 on an outlet at k = mz-1, and
// ibm_interpolation_advanced when the context is immersed and a body is
// given (rhs.c:1604-1614).  Then the four stages, each Contra2Cart;
// Rhs = Formfunction_2(U) - dP; U = alfa dt st Rhs + U0; Contra2Cart
// (rhs.c:1637-1659), with U0 the flux after the boundary work.  After the
// stages: FormBCS's outlet branch with outflow_scale = 1, which spreads the
// deficit between the inflow and the outflow over the outlet faces
// (rhs.c:1685-1686), so the Poisson right-hand side sees a balanced flux.
#include "vfs_internal.h"

namespace {

// rhs -= dP (and += F when a body force is carried); ucont = a*rhs + u0
// (rhs.c:1646-1650)
__global__ void k_rk_update_f(long long n, double a, const double *__restrict__ dP, const double *__restrict__ f,
                              const double *__restrict__ u0, double *__restrict__ rhs, double *__restrict__ ucont)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double r = rhs[t];
    r += -1 * dP[t];
    if (f) r += 1. * f[t];
    rhs[t] = r;
    ucont[t] = a * r + u0[t];
}

__global__ void k_copy(long long n, const double *__restrict__ a, double *__restrict__ b)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

}   // namespace

extern "C" int vfs_runge_kutta(vfs_ctx *c, const vfs_rk_bc *bc, double *ucont, double *ucat, const double *nu_t,
                               const double *dP, double *u0, double *rhs, double *flux_in, double *flux_out)
{
    const Dims d = c->d;
    VFS_REQUIRE(ucont && ucat && dP && u0 && rhs, "bad arguments");
    const bool inlet = !d.kp && d.bc[4] == 5, outlet = !d.kp && d.bc[5] == 4;
    VFS_REQUIRE(!(inlet || outlet) || (bc && bc->ubcs), "an open k end needs bc->ubcs");
    VFS_REQUIRE(!inlet || (bc->inflow && bc->coor), "an inlet needs bc->inflow and bc->coor");
    VFS_REQUIRE(!(bc && bc->surf) || (bc->nodes && bc->p && bc->ustar && bc->n_nodes >= 0),
                "a body needs its node list, p and ustar");
    const double *force = bc && bc->add_body_force ? c->body_force : nullptr;
    VFS_REQUIRE(!(bc && bc->add_body_force) || force, "add_body_force without a body force (vfs_set_body_force)");
    double fin = 0, fout = 0, area = 0;
    // rhs.c:1604-1614
    if (inlet && vfs_inflow_flux(c, bc->inflow, bc->coor, bc->uin_plane, ucont, ucat, bc->ubcs, &fin)) return -1;
    if (outlet && vfs_outlet_correction(c, fin, 0, ucat, ucont, bc->ubcs, &fout, &area)) return -1;
    if (d.immersed && bc && bc->surf) {
        const vfs_ib_options dflt = {0, 0, 0.0, 0.0};
        if (vfs_ib_interpolate_opt(c, bc->surf, bc->nodes, bc->n_nodes, bc->ib ? bc->ib : &dflt, ucont, ucat, bc->p,
                                   bc->ustar, nullptr, nullptr, nullptr))
            return -1;
    }
    const long long n3 = 3 * d.P;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n3, 256)), dim3(256), 0, c->stream, n3, ucont, u0);
    VFS_CHECK_LAUNCH();
    // rhs.c:1633-1662
    const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
    for (int s = 0; s < 4; s++) {
        if (vfs_contra2cart(c, ucont, ucat)) return -1;
        VFS_CHECK_HIP(hipMemsetAsync(rhs, 0, (size_t)n3 * sizeof(double), c->stream));
        if (vfs_formfunction2(c, ucont, ucat, nu_t, rhs, 1.0)) return -1;
        const double a = alfa[s] * d.dt * d.st;
        hipLaunchKernelGGL(k_rk_update_f, dim3(grid_for(n3, 256)), dim3(256), 0, c->stream, n3, a, dP, force, u0,
                           rhs, ucont);
        VFS_CHECK_LAUNCH();
        if (vfs_contra2cart(c, ucont, ucat)) return -1;
    }
    // rhs.c:1685-1686
    if (outlet && vfs_outlet_correction(c, fin, 1, ucat, ucont, bc->ubcs, &fout, &area)) return -1;
    if (flux_in) *flux_in = fin;
    if (flux_out) *flux_out = fout;
    return 0;
}
