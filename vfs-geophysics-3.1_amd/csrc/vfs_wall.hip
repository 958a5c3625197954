// Wall-function walls (bctype -1 on an i or j wall): the pieces of the
// reference that implement them around the hot path.
//  - the IB-node mask of the first wall cells when no body is immersed
//    (IB_BC at the first step, momentum.c:2014-2039);
//  - the wall-model velocity and friction velocity of those cells in
//    Contra2Cart_2 (rhs.c:303-333, 369-401; wallfunction.c:18-43);
//  - the zero wall-face flux of IB_BC (momentum.c:2135-2153).
// The rest of the path sees these walls through the mask.
#include <vector>

#include "vfs_internal.h"
#include "vfs_wall.h"

// wallfunction.c:291-323 pre_integrate, on the host (same libm exp as the
// reference, same operation order): the cumulative integral of
// 1 / (1 + nu_t/nu) over y+ in steps of 2, each by 24 Simpson 3/8 panels.
static void wf_build_table(std::vector<double> &b)
{
    b.assign(VFS_WF_NYP + 1, 0.0);
    b[0] = 0.;
    double E[25];
    for (int i = 1; i <= VFS_WF_NYP; i++) {
        const int N = 24;
        const double ya_p = (double)(i - 1) * VFS_WF_INTERVAL;
        const double yb_p = (double)(i)*VFS_WF_INTERVAL;
        const double ydiff = yb_p - ya_p, dy_p = ydiff / (double)N;
        double val = 0, ybegin_p = ya_p;
        for (int k = 0; k <= N; k++) E[k] = 1. / (1. + wf_nut(ybegin_p + dy_p * k));
        for (int k = 0; k < N; k++) {
            double F[4];
            F[0] = E[k];
            F[1] = 1. / (1. + wf_nut(ybegin_p + dy_p * 1. / 3.));
            F[2] = 1. / (1. + wf_nut(ybegin_p + dy_p * 2. / 3.));
            F[3] = E[k + 1];
            val += dy_p / 3. * (3 * F[0] + 9 * F[1] + 9 * F[2] + 3 * F[3]) / 8.;
            ybegin_p += dy_p;
        }
        b[i] = b[i - 1] + val;
    }
}

bool vfs_has_wallfn(const Dims &d)
{
    for (int f = 0; f < 4; f++)
        if (d.bc[f] == -1) return true;
    return false;
}

// table and friction-velocity field of a context with wall-function walls
int vfs_wall_setup(vfs_ctx *c)
{
    if (!vfs_has_wallfn(c->d) || c->wf_table) return 0;
    std::vector<double> b;
    wf_build_table(b);
    VFS_CHECK_HIP(hipMalloc(&c->wf_table, b.size() * sizeof(double)));
    VFS_CHECK_HIP(hipMemcpy(c->wf_table, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
    VFS_CHECK_HIP(hipMalloc(&c->ustar, (size_t)c->d.P * sizeof(double)));
    VFS_CHECK_HIP(hipMemset(c->ustar, 0, (size_t)c->d.P * sizeof(double)));
    return 0;
}

// momentum.c:2014-2039 (no immersed body): nvert = 1 on the first cells of
// the wall-function walls
__global__ void k_wall_mask(Dims d, double *nvert)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    if ((d.bc[0] == -1 && i == 1) || (d.bc[1] == -1 && i == d.mx - 2) || (d.bc[2] == -1 && j == 1) ||
        (d.bc[3] == -1 && j == d.my - 2) || (d.bc[4] == -1 && k == 1) || (d.bc[5] == -1 && k == d.mz - 2))
        nvert[t] = 1;
}

int vfs_wall_mask(vfs_ctx *c)
{
    c->nvert_version++;
    const Dims d = c->d;
    if (d.immersed || !vfs_has_wallfn(d)) return 0;
    hipLaunchKernelGGL(k_wall_mask, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->nvert);
    VFS_CHECK_LAUNCH();
    return 0;
}

// Contra2Cart_2's wall-function branch (rhs.c:303-333 for i walls, 369-401
// for j walls) on the first cells of wall walls of direction dir: sb and sc
// are the wall distances of this cell's and the next cell's centres, Uc the
// next cell's velocity from the snapshot L.  One thread per wall point;
// grid (ceil(n1/64), mz-2, 2) over the transverse index (j for i walls, i for
// j walls), k and the two walls.
__global__ void k_c2c_wall(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ tab,
                           const double *__restrict__ L, int dir, double *__restrict__ ucat, double *__restrict__ ustar)
{
    const int t1 = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y, side = blockIdx.z;
    const int m = dir == 0 ? d.mx : d.my, m1 = dir == 0 ? d.my : d.mx;
    if (t1 > m1 - 2 || d.bc[2 * dir + side] != -1) return;
    const int n = side == 0 ? 1 : m - 2;
    const int i = dir == 0 ? n : t1, j = dir == 0 ? t1 : n;
    const long long id = IX(d, k, j, i);
    if ((int)(nvert[id] + 0.1) == 3) return;   // solid: zeroed by the boundary loop (rhs.c:296-300)
    const double *a = (dir == 0 ? g.csi : g.eta) + 3 * id;
    const double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double sb = 0.5 / g.aj[id] / area;
    const long long st = dir == 0 ? 1 : (long long)d.mx;
    const long long nb = n == 1 ? id + st : id - st;
    const double sc = 2 * sb + 0.5 / g.aj[nb] / area;
    const double Uc[3] = {L[3 * nb], L[3 * nb + 1], L[3 * nb + 2]};
    double nrm[3], Ub[3], us;
    wf_normal(g.csi + 3 * id, g.eta + 3 * id, g.zet + 3 * id, dir, nrm);
    if (n == m - 2) nrm[0] *= -1, nrm[1] *= -1, nrm[2] *= -1;
    wf_wall_function(tab, 1. / d.ren, sc, sb, Uc, Ub, &us, nrm[0], nrm[1], nrm[2]);
    ustar[id] = us;
    ucat[3 * id] = Ub[0];
    ucat[3 * id + 1] = Ub[1];
    ucat[3 * id + 2] = Ub[2];
}

// i walls first, then j walls (a corner cell of both keeps the j result, as
// in the reference's loop order)
int vfs_c2c_wall(vfs_ctx *c, const double *L, double *ucat)
{
    const Dims d = c->d;
    if (!vfs_has_wallfn(d)) return 0;
    if (vfs_wall_setup(c)) return -1;
    for (int dir = 0; dir < 2; dir++) {
        if (d.bc[2 * dir] != -1 && d.bc[2 * dir + 1] != -1) continue;
        const int n1 = (dir == 0 ? d.my : d.mx) - 2;
        hipLaunchKernelGGL(k_c2c_wall, dim3((n1 + 63) / 64, d.mz - 2, 2), dim3(64), 0, c->stream, d, c->g, c->nvert,
                           c->wf_table, L, dir, ucat, c->ustar);
        VFS_CHECK_LAUNCH();
    }
    return 0;
}

// IB_BC's wall-function faces (momentum.c:2135-2153): the flux through the
// wall face of each first cell is zero.  grid (ceil(n1/64), mz-2, 2)
__global__ void k_ibbc_wall(Dims d, int dir, double *__restrict__ ucont)
{
    const int t1 = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y, side = blockIdx.z;
    const int m = dir == 0 ? d.mx : d.my, m1 = dir == 0 ? d.my : d.mx;
    if (t1 > m1 - 2 || d.bc[2 * dir + side] != -1) return;
    const int n = side == 0 ? 0 : m - 2;   // the face below cell 1, or above cell m-2
    const int i = dir == 0 ? n : t1, j = dir == 0 ? t1 : n;
    ucont[3 * IX(d, k, j, i) + dir] = 0;
}

int vfs_ibbc_wall(vfs_ctx *c, double *ucont)
{
    const Dims d = c->d;
    for (int dir = 0; dir < 2; dir++) {
        if (d.bc[2 * dir] != -1 && d.bc[2 * dir + 1] != -1) continue;
        const int n1 = (dir == 0 ? d.my : d.mx) - 2;
        hipLaunchKernelGGL(k_ibbc_wall, dim3((n1 + 63) / 64, d.mz - 2, 2), dim3(64), 0, c->stream, d, dir, ucont);
        VFS_CHECK_LAUNCH();
    }
    return 0;
}

extern "C" int vfs_get_ustar(vfs_ctx *c, double *dst)
{
    VFS_REQUIRE(vfs_has_wallfn(c->d), "no wall-function wall (bctype -1) in this context");
    if (vfs_wall_setup(c)) return -1;
    VFS_CHECK_HIP(hipMemcpyAsync(dst, c->ustar, (size_t)c->d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}
