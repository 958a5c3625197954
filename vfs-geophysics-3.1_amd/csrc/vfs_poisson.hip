// Pressure projection kernels on gfx950: setup_lidx2 (poisson.c:182-329),
// PoissonLHSNew's 19-point curvilinear stencil (poisson.c:971-1682),
// PoissonRHS2_hypre + Remove_Nullspace_Scale (poisson_hypre.c:110-399),
// VolumeFlux (poisson.c:3570-3910), UpdatePressure (3144-3304), Projection
// (2707-3073) and Divergence (main.c:1047-1177).
//
// The stencil is assembled once per mask into slot-major SoA arrays
// (coef[m*P + cell], m in the reference's slot order CP..BW), so the
// operator apply streams 19 coalesced coefficient planes plus x and y:
// 168 B/cell in FP64, 84 B/cell in FP32.
#include "vfs_internal.h"
#include "vfs_poisson.h"
#include "vfs_reduce.h"

#define NVv(kk, jj, ii) nvert[IX(d, kk, jj, ii)]

// neighbour offsets of the 19 slots (di, dj, dk), poisson.c:951-969 and 1562-1580
__constant__ int c_slot[19][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1},
    {0, -1, -1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1}};

enum { CP, EP, WP, NP, SP, TP, BP, NE, SE, NW, SW, TN, BN, TS, BS, TE, BE, TW, BW };

// g_ab of poisson.c:1096-1132 with the periodic index map
__device__ __forceinline__ double gfield(const Dims &d, const Geom &g, int which, int k, int j, int i)
{
    int a, b, c;
    per_src(d, i, j, k, a, b, c);
    long long s = IX(d, c, b, a);
    const double *x, *y, *w;
    switch (which) {
    case 11: x = g.icsi; y = g.icsi; w = g.iaj; break;
    case 12: x = g.ieta; y = g.icsi; w = g.iaj; break;
    case 13: x = g.izet; y = g.icsi; w = g.iaj; break;
    case 21: x = g.jcsi; y = g.jeta; w = g.jaj; break;
    case 22: x = g.jeta; y = g.jeta; w = g.jaj; break;
    case 23: x = g.jzet; y = g.jeta; w = g.jaj; break;
    case 31: x = g.kcsi; y = g.kzet; w = g.kaj; break;
    case 32: x = g.keta; y = g.kzet; w = g.kaj; break;
    default: x = g.kzet; y = g.kzet; w = g.kaj; break;
    }
    return (x[3 * s] * y[3 * s] + x[3 * s + 1] * y[3 * s + 1] + x[3 * s + 2] * y[3 * s + 2]) * w[s];
}

// unknown test of setup_lidx2 (poisson.c:231-250) through the periodic map
__device__ __forceinline__ bool is_unknown(const Dims &d, const double *nvert, int k, int j, int i)
{
    int a, b, c;
    per_src(d, i, j, k, a, b, c);
    if (a < 1 || b < 1 || c < 1 || a > d.mx - 2 || b > d.my - 2 || c > d.mz - 2) return false;
    return !((int)nvert[IX(d, c, b, a)] > 0.1);
}

// One fluid row of PoissonLHSNew (poisson.c:1160-1555), literal order.
__device__ __forceinline__ void stencil_row(const Dims &d, const Geom &g, const double *__restrict__ nvert, int i,
                                            int j, int k, double vol[19])
{
    const int ip = d.ip, jp = d.jp, kp = d.kp, mx = d.mx, my = d.my, mz = d.mz;
    const double th = 0.1, r = 1.0;
#define NV NVv
#define G(w, kk, jj, ii) gfield(d, g, w, kk, jj, ii)
#include "vfs_stencil_row.inc"
#undef G
#undef NV
}

__global__ void k_poisson_coef(Dims d, Geom g, const double *__restrict__ nvert, double *__restrict__ coef,
                               float *__restrict__ coef32, unsigned char *__restrict__ rowflag)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double vol[19];
#pragma unroll
    for (int m = 0; m < 19; m++) vol[m] = 0.;
    unsigned char flag = 0;   // 0: no row, 1: fluid row, 2: identity row
    bool interior = i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2;
    if (interior) {
        int row_ok = !((int)nvert[t] > 0.1);
        if (nvert[t] > 0.1 && row_ok) { vol[CP] = 1.0; flag = 2; }
        else if (!(nvert[t] > 0.1)) {
            stencil_row(d, g, nvert, i, j, k, vol);
            flag = 1;
            // drop rule (poisson.c:1604-1608); unrolled, so the row stays in registers
#pragma unroll
            for (int m = 1; m < 19; m++) {
                bool keep = fabs(vol[m]) > 1.e-10 &&
                            is_unknown(d, nvert, k + c_slot[m][2], j + c_slot[m][1], i + c_slot[m][0]);
                if (!keep) vol[m] = 0.;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 19; m++) {
        coef[(long long)m * d.P + t] = vol[m];
        coef32[(long long)m * d.P + t] = (float)vol[m];
    }
    rowflag[t] = flag;
}

// ---------------------------------------------------------------------------
// unknown numbering: exclusive scan of the unknown flags in DA order
// ---------------------------------------------------------------------------
__global__ void k_gid_count(Dims d, const double *__restrict__ nvert, int *blocksum)
{
    __shared__ double sh[4];
    long long t = (long long)blockIdx.x * 1024 + threadIdx.x * 4;
    int cnt = 0;
    for (int q = 0; q < 4; q++) {
        long long u = t + q;
        if (u >= d.P) break;
        int i = (int)(u % d.mx), j = (int)((u / d.mx) % d.my), k = (int)(u / ((long long)d.mx * d.my));
        if (i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2 && !((int)nvert[u] > 0.1)) cnt++;
    }
    double s = block_sum<256>((double)cnt, sh);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = (int)s;
}

__global__ void k_gid_scan_blocks(int nb, int *blocksum, int *total)
{
    // single workgroup: sequential chunks of 256 with a LDS scan
    __shared__ int sh[256];
    int carry = 0;
    for (int base = 0; base < nb; base += 256) {
        int v = base + threadIdx.x < nb ? blocksum[base + threadIdx.x] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            int add = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        int incl = sh[threadIdx.x];
        if (base + threadIdx.x < nb) blocksum[base + threadIdx.x] = carry + incl - v;
        __syncthreads();
        carry += sh[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_gid_fill(Dims d, const double *__restrict__ nvert, const int *__restrict__ blockoff, double *gid)
{
    __shared__ int sh[256];
    long long t = (long long)blockIdx.x * 1024 + threadIdx.x * 4;
    int f[4], cnt = 0;
    for (int q = 0; q < 4; q++) {
        long long u = t + q;
        f[q] = 0;
        if (u < d.P) {
            int i = (int)(u % d.mx), j = (int)((u / d.mx) % d.my), k = (int)(u / ((long long)d.mx * d.my));
            f[q] = i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2 && !((int)nvert[u] > 0.1);
        }
        cnt += f[q];
    }
    sh[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int add = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int run = blockoff[blockIdx.x] + sh[threadIdx.x] - cnt;
    for (int q = 0; q < 4; q++) {
        long long u = t + q;
        if (u >= d.P) break;
        gid[u] = f[q] ? (double)(run++) : -1.0;
    }
}

// ---------------------------------------------------------------------------
// operator apply on the DA layout
// ---------------------------------------------------------------------------
__global__ void k_poisson_apply(Dims d, const double *__restrict__ coef, const unsigned char *__restrict__ rowflag,
                                const double *__restrict__ x, double *__restrict__ y)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    if (!rowflag[t]) { y[t] = 0; return; }
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double s = 0;
    for (int m = 0; m < 19; m++) {
        double c = coef[(long long)m * d.P + t];
        if (c == 0.0) continue;
        int ii = i + c_slot[m][0], jj = j + c_slot[m][1], kk = k + c_slot[m][2];
        int a, b, cc;
        per_src(d, ii, jj, kk, a, b, cc);
        s += c * x[IX(d, cc, b, a)];
    }
    y[t] = s;
}

// ---------------------------------------------------------------------------
// setup / accessors
// ---------------------------------------------------------------------------
int vfs_poisson_alloc(vfs_ctx *c);
int vfs_mg_setup(vfs_ctx *c);

extern "C" int vfs_poisson_setup(vfs_ctx *c, int *n_unknowns)
{
    const Dims d = c->d;
    if (vfs_poisson_alloc(c)) return -1;
    Poisson *ps = c->pois;
    int nb = (int)grid_for(d.P, 1024);
    hipLaunchKernelGGL(k_gid_count, dim3(nb), dim3(256), 0, c->stream, d, c->nvert, ps->blocksum);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gid_scan_blocks, dim3(1), dim3(256), 0, c->stream, nb, ps->blocksum, ps->total);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gid_fill, dim3(nb), dim3(256), 0, c->stream, d, c->nvert, ps->blocksum, ps->gid);
    VFS_CHECK_LAUNCH();
    if (vfs_periodic_sca(c, ps->gid)) return -1;
    hipLaunchKernelGGL(k_poisson_coef, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert,
                       ps->coef, ps->coef32, ps->rowflag);
    VFS_CHECK_LAUNCH();
    int total = 0;
    VFS_CHECK_HIP(hipMemcpyAsync(&total, ps->total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    ps->n = total;
    if (n_unknowns) *n_unknowns = total;
    if (total == 0) {
        vfs_set_error("vfs_poisson_setup: no fluid cell (every cell has nvert >= 0.1), the pressure system is empty");
        return -8;
    }
    if (vfs_mg_setup(c)) return -1;
    ps->ready = 1;
    return 0;
}

extern "C" int vfs_poisson_gid(vfs_ctx *c, double *gid)
{
    VFS_REQUIRE(c->pois && c->pois->ready, "vfs_poisson_setup has not been called");
    VFS_CHECK_HIP(hipMemcpyAsync(gid, c->pois->gid, (size_t)c->d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

extern "C" int vfs_poisson_coefficients(vfs_ctx *c, double *coef)
{
    VFS_REQUIRE(c->pois && c->pois->ready, "vfs_poisson_setup has not been called");
    VFS_CHECK_HIP(hipMemcpyAsync(coef, c->pois->coef, 19 * (size_t)c->d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

extern "C" int vfs_poisson_apply(vfs_ctx *c, const double *x, double *y)
{
    VFS_REQUIRE(c->pois && c->pois->ready, "vfs_poisson_setup has not been called");
    const Dims d = c->d;
    hipLaunchKernelGGL(k_poisson_apply, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->pois->coef,
                       c->pois->rowflag, x, y);
    VFS_CHECK_LAUNCH();
    return 0;
}

// ---------------------------------------------------------------------------
// Poisson right-hand side
// ---------------------------------------------------------------------------
__global__ void k_prhs(Dims d, const double *__restrict__ nvert, const double *__restrict__ aj,
                       const double *__restrict__ ucont, double *__restrict__ b, double *part)
{
    __shared__ double sh[4];
    double sb = 0, sa = 0;
    const double coeff = 1.0;   // time_coeff(), poisson.c:17-25
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < d.P; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
        double v = 0;
        if (i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2 && !(nvert[t] >= 0.1)) {
            v -= ucont[3 * t];
            if (i == 1 && d.ip) v += ucont[3 * IX(d, k, j, d.mx - 2)];
            else v += ucont[3 * (t - 1)];
            v -= ucont[3 * t + 1];
            if (j == 1 && d.jp) v += ucont[3 * IX(d, k, d.my - 2, i) + 1];
            else v += ucont[3 * (t - d.mx) + 1];
            v -= ucont[3 * t + 2];
            if (k == 1 && d.kp) v += ucont[3 * IX(d, d.mz - 2, j, i) + 2];
            else v += ucont[3 * (t - (long long)d.mx * d.my) + 2];
            v *= -1.0 / d.dt * d.st * coeff;
            if (nvert[t] < 0.1) { sb += v; sa += aj[t]; }
        }
        b[t] = v;
    }
    double r0 = block_sum<256>(sb, sh);
    double r1 = block_sum<256>(sa, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = r0; part[VFS_RED_BLOCKS + blockIdx.x] = r1; }
}

__global__ void k_prhs_scale(Dims d, const double *__restrict__ nvert, const double *__restrict__ aj,
                             const double *__restrict__ sums, double *__restrict__ b)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    if (nvert[t] < 0.1) b[t] += -sums[0] * aj[t] / sums[1];
}

extern "C" int vfs_poisson_rhs(vfs_ctx *c, const double *ucont, double *b)
{
    const Dims d = c->d;
    int nb = (int)grid_for(d.P, 256);
    if (nb > VFS_RED_BLOCKS) nb = VFS_RED_BLOCKS;
    hipLaunchKernelGGL(k_prhs, dim3(nb), dim3(256), 0, c->stream, d, c->nvert, c->g.aj, ucont, b, c->red);
    VFS_CHECK_LAUNCH();
    double *res = c->red + VFS_RED_SLOTS * VFS_RED_BLOCKS;
    if (vfs_reduce_finish(c, nb, 2, 0, res)) return -1;
    hipLaunchKernelGGL(k_prhs_scale, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->nvert, c->g.aj, res, b);
    VFS_CHECK_LAUNCH();
    return 0;
}

// ---------------------------------------------------------------------------
// VolumeFlux (immersed==1, flg=0): pass 0 zero small IB fluxes + sum,
// pass 1 correct, pass 2 sum again
// ---------------------------------------------------------------------------
// VolumeFlux, immersed == 1 (poisson.c:3570-3910) or, with plain != 0, the
// variant for immersed == 0 (poisson.c:3306-3500): every fluid / non-fluid
// face (no upper nvert bound, last face included, no small-flux zeroing),
// whose fluxes the correction pass sets to zero.
__global__ void k_vflux(Dims d, Geom g, const double *__restrict__ nvert, double *__restrict__ ucont, int pass,
                        const double *__restrict__ sums, double *part, int plain)
{
    __shared__ double sh[4];
    const double eps = 1.e-8, ibmval = 1.1;
    double corr = 0;
    if (pass == 1) corr = sums[1] > 1.e-15 ? sums[0] / sums[1] : 0;
    double fl = 0, ar = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < d.P; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
        if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) continue;
        const double *arr[3] = {g.csi, g.eta, g.zet};
        const double nv = nvert[t];
        const double nb[3] = {nvert[t + 1], nvert[t + d.mx], nvert[t + (long long)d.mx * d.my]};
        const bool inb[3] = {i < d.mx - 2, j < d.my - 2, k < d.mz - 2};
        for (int side = 0; side < 2; side++) {
            for (int q = 0; q < 3; q++) {
                bool hit;
                if (plain) hit = side == 0 ? (nv < 0.1 && nb[q] > 0.1) : (nv > 0.1 && nb[q] < 0.1);
                else hit = side == 0 ? (nv < 0.1 && nb[q] > 0.1 && nb[q] < ibmval && inb[q])
                                     : (nv > 0.1 && nv < ibmval && nb[q] < 0.1 && inb[q]);
                if (!hit) continue;
                const double *av = arr[q] + 3 * t;
                double a = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
                double u = ucont[3 * t + q];
                double sgn = side == 0 ? 1.0 : -1.0;
                if (plain) {
                    if (pass == 1) ucont[3 * t + q] = 0.;   // corrected, then zeroed (poisson.c:3400-3430)
                    else { fl = side == 0 ? fl + u : fl - u; ar += a; }
                    continue;
                }
                if (pass == 0) {
                    if (fabs(u) > eps) { fl = side == 0 ? fl + u : fl - u; ar += a; }
                    else ucont[3 * t + q] = 0.;
                } else if (pass == 1) {
                    if (fabs(u) > eps) ucont[3 * t + q] = side == 0 ? u - a * corr : u + a * corr;
                } else {
                    fl = side == 0 ? fl + u : fl - u;
                    ar += a;
                }
                (void)sgn;
            }
        }
    }
    if (pass != 1) {
        double r0 = block_sum<256>(fl, sh);
        double r1 = block_sum<256>(ar, sh);
        if (threadIdx.x == 0) { part[blockIdx.x] = r0; part[VFS_RED_BLOCKS + blockIdx.x] = r1; }
    }
}

extern "C" int vfs_volume_flux(vfs_ctx *c, double *ucont, double *flux, double *area)
{
    const Dims d = c->d;
    double *res = c->red + VFS_RED_SLOTS * VFS_RED_BLOCKS;
    const int plain = !d.immersed;
    int nb = (int)grid_for(d.P, 256);
    if (nb > VFS_RED_BLOCKS) nb = VFS_RED_BLOCKS;
    hipLaunchKernelGGL(k_vflux, dim3(nb), dim3(256), 0, c->stream, d, c->g, c->nvert, ucont, 0, res, c->red, plain);
    VFS_CHECK_LAUNCH();
    if (vfs_reduce_finish(c, nb, 2, 0, res)) return -1;
    hipLaunchKernelGGL(k_vflux, dim3(nb), dim3(256), 0, c->stream, d, c->g, c->nvert, ucont, 1, res, c->red, plain);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vflux, dim3(nb), dim3(256), 0, c->stream, d, c->g, c->nvert, ucont, 2, res, c->red, plain);
    VFS_CHECK_LAUNCH();
    if (vfs_reduce_finish(c, nb, 2, 0, res + 2)) return -1;
    if (flux || area) {
        double h[2];
        if (vfs_read_scalars(c, res + 2, h, 2)) return -1;
        if (flux) *flux = h[0];
        if (area) *area = h[1];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// UpdatePressure, Projection, Divergence
// ---------------------------------------------------------------------------
__global__ void k_update_p(Dims d, const double *__restrict__ nvert, const double *__restrict__ phi, double *__restrict__ p)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    double v = p[t];
    if (nvert[t] < 0.1) v += phi[t];
    if (nvert[t] > 1.1) v = 0;
    p[t] = v;
}

extern "C" int vfs_update_pressure(vfs_ctx *c, double *p, double *phi)
{
    const Dims d = c->d;
    hipLaunchKernelGGL(k_update_p, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->nvert, phi, p);
    VFS_CHECK_LAUNCH();
    if (vfs_periodic_sca(c, p)) return -1;
    return vfs_periodic_sca(c, phi);
}

__global__ void k_projection(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ phi,
                             double *__restrict__ ucont)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
#define NV NVv
#define PH(kk, jj, ii) phi[IX(d, kk, jj, ii)]
#include "vfs_projection.inc"
#undef PH
#undef NV
}

__global__ void k_projection_periodic(Dims d, double *ucont)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    int fi = 0, fj = 0, fk = 0, a = i, b = j, cc = k;
    if (d.ip && i == 0) a = d.mx - 2, fi = 1; else if (d.ip && i == d.mx - 1) a = 1, fi = 1;
    if (d.jp && j == 0) b = d.my - 2, fj = 1; else if (d.jp && j == d.my - 1) b = 1, fj = 1;
    if (d.kp && k == 0) cc = d.mz - 2, fk = 1; else if (d.kp && k == d.mz - 1) cc = 1, fk = 1;
    long long o = IX(d, k, j, i), s = IX(d, cc, b, a);
    if (fi) ucont[3 * o] = ucont[3 * s];
    if (fj) ucont[3 * o + 1] = ucont[3 * s + 1];
    if (fk) ucont[3 * o + 2] = ucont[3 * s + 2];
}

extern "C" int vfs_projection(vfs_ctx *c, const double *phi, double *ucont)
{
    const Dims d = c->d;
    hipLaunchKernelGGL(k_projection, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, phi, ucont);
    VFS_CHECK_LAUNCH();
    if (d.ip || d.jp || d.kp) {
        hipLaunchKernelGGL(k_projection_periodic, dim3(grid_for(shell_count(d), 256)), dim3(256), 0, c->stream, d, ucont);
        VFS_CHECK_LAUNCH();
    }
    return 0;
}

__global__ void k_divergence(Dims d, const double *__restrict__ nvert, const double *__restrict__ aj,
                             const double *__restrict__ ucont, double *div, double *part)
{
    __shared__ double sh[4];
    double mxv = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < d.P; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
        double m = 0;
        if (i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2) {
            const long long sj = d.mx, sk = (long long)d.mx * d.my;
            m = fabs((ucont[3 * t] - ucont[3 * (t - 1)] + ucont[3 * t + 1] - ucont[3 * (t - sj) + 1] +
                      ucont[3 * t + 2] - ucont[3 * (t - sk) + 2]) * aj[t]);
            if (nvert[t] + nvert[t + sk] + nvert[t - sk] + nvert[t + sj] + nvert[t - sj] + nvert[t + 1] + nvert[t - 1] > 0.1) m = 0.;
        }
        if (div) div[t] = m;
        mxv = fmax(mxv, m);
    }
    double r = block_max<256>(mxv, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

extern "C" int vfs_divergence(vfs_ctx *c, const double *ucont, double *div, double *maxdiv)
{
    const Dims d = c->d;
    int nb = (int)grid_for(d.P, 256);
    if (nb > VFS_RED_BLOCKS) nb = VFS_RED_BLOCKS;
    hipLaunchKernelGGL(k_divergence, dim3(nb), dim3(256), 0, c->stream, d, c->nvert, c->g.aj, ucont, div, c->red);
    VFS_CHECK_LAUNCH();
    double *res = c->red + VFS_RED_SLOTS * VFS_RED_BLOCKS + 8;
    if (vfs_reduce_finish(c, nb, 1, 1, res)) return -1;
    if (maxdiv) return vfs_read_scalars(c, res, maxdiv, 1);
    return 0;
}
