// FormMetrics (metrics.c:4-1346, NEWMETRIC) and the immersed-boundary mask
// (ibm.c:224, 319-341, 508-516) as gfx950 kernels.
//
// Set-up work (once per grid): one thread per point, i fastest, so every
// wavefront reads 64 consecutive nodes of a grid line.  The sequence of
// passes reproduces the reference's loop order where later passes overwrite
// earlier ones (boundary mirror, then periodic copies, then face averages,
// then the final periodic copy).
#include "vfs_internal.h"
#include "vfs_reduce.h"

#define CO(k, j, i, q) coor[3 * IX(d, k, j, i) + (q)]

// cell centres, Jacobian and csi/eta/zet of interior cells (metrics.c:80-99, 186-251)
__global__ void k_cell_metrics(Dims d, const double *__restrict__ coor, Geom g)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    double dd[3][3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        g.cent[3 * t + q] = 0.125 * (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k - 1, j, i, q) + CO(k - 1, j - 1, i, q) +
                                     CO(k, j, i - 1, q) + CO(k, j - 1, i - 1, q) + CO(k - 1, j, i - 1, q) + CO(k - 1, j - 1, i - 1, q));
        dd[q][0] = 0.25 * (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k - 1, j, i, q) + CO(k - 1, j - 1, i, q) -
                           CO(k, j, i - 1, q) - CO(k, j - 1, i - 1, q) - CO(k - 1, j, i - 1, q) - CO(k - 1, j - 1, i - 1, q));
        dd[q][1] = 0.25 * (CO(k, j, i, q) + CO(k, j, i - 1, q) + CO(k - 1, j, i, q) + CO(k - 1, j, i - 1, q) -
                           CO(k, j - 1, i, q) - CO(k, j - 1, i - 1, q) - CO(k - 1, j - 1, i, q) - CO(k - 1, j - 1, i - 1, q));
        dd[q][2] = 0.25 * (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k, j, i - 1, q) + CO(k, j - 1, i - 1, q) -
                           CO(k - 1, j, i, q) - CO(k - 1, j - 1, i, q) - CO(k - 1, j, i - 1, q) - CO(k - 1, j - 1, i - 1, q));
    }
    double dxdc = dd[0][0], dydc = dd[1][0], dzdc = dd[2][0];
    double dxde = dd[0][1], dyde = dd[1][1], dzde = dd[2][1];
    double dxdz = dd[0][2], dydz = dd[1][2], dzdz = dd[2][2];
    double a = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) + dzdc * (dxde * dydz - dyde * dxdz);
    g.aj[t] = 1. / a;
    g.csi[3 * t] = dyde * dzdz - dzde * dydz;
    g.csi[3 * t + 1] = -dxde * dzdz + dzde * dxdz;
    g.csi[3 * t + 2] = dxde * dydz - dyde * dxdz;
    g.eta[3 * t] = dydz * dzdc - dzdz * dydc;
    g.eta[3 * t + 1] = -dxdz * dzdc + dzdz * dxdc;
    g.eta[3 * t + 2] = dxdz * dydc - dydz * dxdc;
    g.zet[3 * t] = dydc * dzde - dzdc * dyde;
    g.zet[3 * t + 1] = -dxdc * dzde + dzdc * dxde;
    g.zet[3 * t + 2] = dxdc * dyde - dydc * dxde;
}

// boundary mirror of one direction (metrics.c:253-334); dir 0: i planes, 1: j, 2: k.
// The two planes of one direction never read each other, so both run at once.
__global__ void k_mirror(Dims d, Geom g, int dir)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int n0 = dir == 0 ? d.my : d.mx, n1 = dir == 2 ? d.my : d.mz;
    long long plane = (long long)n0 * n1;
    if (t >= 2 * plane) return;
    int side = (int)(t / plane);
    long long r = t % plane;
    int a = (int)(r % n0), b = (int)(r / n0);
    int i, j, k, si, sj, sk;
    if (dir == 0) { j = a; k = b; i = side ? d.mx - 1 : 0; si = side ? d.mx - 2 : 1; sj = j; sk = k; }
    else if (dir == 1) { i = a; k = b; j = side ? d.my - 1 : 0; sj = side ? d.my - 2 : 1; si = i; sk = k; }
    else { i = a; j = b; k = side ? d.mz - 1 : 0; sk = side ? d.mz - 2 : 1; si = i; sj = j; }
    long long o = IX(d, k, j, i), s = IX(d, sk, sj, si);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        g.csi[3 * o + q] = g.csi[3 * s + q];
        g.eta[3 * o + q] = g.eta[3 * s + q];
        g.zet[3 * o + q] = g.zet[3 * s + q];
    }
    g.aj[o] = g.aj[s];
}

// periodic ghosts of the cell metrics (metrics.c:354-389)
__global__ void k_periodic_cellmetrics(Dims d, Geom g)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i, a, b, c;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    if (!per_src(d, i, j, k, a, b, c)) return;
    long long o = IX(d, k, j, i), s = IX(d, c, b, a);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        g.csi[3 * o + q] = g.csi[3 * s + q];
        g.eta[3 * o + q] = g.eta[3 * s + q];
        g.zet[3 * o + q] = g.zet[3 * s + q];
    }
    g.aj[o] = g.aj[s];
}

// face metrics: averages of the two adjacent cells and the harmonic-mean
// Jacobian (metrics.c:581-584, 689-692, 797-800)
__global__ void k_face_metrics(Dims d, Geom g)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (k >= 1 && k <= d.mz - 2 && j >= 1 && j <= d.my - 2 && i <= d.mx - 2) {
        long long p = IX(d, k, j, i + 1);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            g.icsi[3 * t + q] = 0.5 * g.csi[3 * t + q] + 0.5 * g.csi[3 * p + q];
            g.ieta[3 * t + q] = 0.5 * g.eta[3 * t + q] + 0.5 * g.eta[3 * p + q];
            g.izet[3 * t + q] = 0.5 * g.zet[3 * t + q] + 0.5 * g.zet[3 * p + q];
        }
        g.iaj[t] = 2. / (1. / g.aj[t] + 1. / g.aj[p]);
    }
    if (k >= 1 && k <= d.mz - 2 && j <= d.my - 2 && i >= 1 && i <= d.mx - 2) {
        long long p = IX(d, k, j + 1, i);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            g.jcsi[3 * t + q] = 0.5 * g.csi[3 * t + q] + 0.5 * g.csi[3 * p + q];
            g.jeta[3 * t + q] = 0.5 * g.eta[3 * t + q] + 0.5 * g.eta[3 * p + q];
            g.jzet[3 * t + q] = 0.5 * g.zet[3 * t + q] + 0.5 * g.zet[3 * p + q];
        }
        g.jaj[t] = 2. / (1. / g.aj[t] + 1. / g.aj[p]);
    }
    if (k <= d.mz - 2 && j >= 1 && j <= d.my - 2 && i >= 1 && i <= d.mx - 2) {
        long long p = IX(d, k + 1, j, i);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            g.kcsi[3 * t + q] = 0.5 * g.csi[3 * t + q] + 0.5 * g.csi[3 * p + q];
            g.keta[3 * t + q] = 0.5 * g.eta[3 * t + q] + 0.5 * g.eta[3 * p + q];
            g.kzet[3 * t + q] = 0.5 * g.zet[3 * t + q] + 0.5 * g.zet[3 * p + q];
        }
        g.kaj[t] = 2. / (1. / g.aj[t] + 1. / g.aj[p]);
    }
}

// final periodic copy of all sixteen arrays (metrics.c:1214-1250)
__global__ void k_periodic_allmetrics(Dims d, Geom g)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i, a, b, c;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    if (!per_src(d, i, j, k, a, b, c)) return;
    long long o = IX(d, k, j, i), s = IX(d, c, b, a);
    double *v[12] = {g.csi, g.eta, g.zet, g.icsi, g.ieta, g.izet, g.jcsi, g.jeta, g.jzet, g.kcsi, g.keta, g.kzet};
    for (int n = 0; n < 12; n++) {
        v[n][3 * o] = v[n][3 * s]; v[n][3 * o + 1] = v[n][3 * s + 1]; v[n][3 * o + 2] = v[n][3 * s + 2];
    }
    g.aj[o] = g.aj[s]; g.iaj[o] = g.iaj[s]; g.jaj[o] = g.jaj[s]; g.kaj[o] = g.kaj[s];
}

// interior cells whose Jacobian is not a positive finite number (the
// reference prints "Negative jacobian" for aj < 0, momentum.c:157): per-block
// count at part[b] and the largest such point index at part[VFS_RED_BLOCKS + b]
__global__ void k_aj_check(Dims d, const double *__restrict__ aj, double *part)
{
    __shared__ double sh[4];
    double cnt = 0, last = -1;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < d.P;
         t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
        if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) continue;
        const double a = aj[t];
        if (!(a > 0.0 && a < INFINITY)) { cnt += 1.0; last = (double)t; }
    }
    double r = block_sum<256>(cnt, sh);
    double m = block_max<256>(last, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = r; part[VFS_RED_BLOCKS + blockIdx.x] = m; }
}

extern "C" int vfs_build_metrics(vfs_ctx *c, const double *coor)
{
    const Dims d = c->d;
    VFS_CHECK_HIP(hipMemsetAsync(c->geom_block, 0, (13 * 3 + 4) * (size_t)d.P * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_cell_metrics, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, coor, c->g);
    VFS_CHECK_LAUNCH();
    for (int dir = 0; dir < 3; dir++) {
        long long n = 2LL * (dir == 0 ? d.my : d.mx) * (dir == 2 ? d.my : d.mz);
        hipLaunchKernelGGL(k_mirror, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d, c->g, dir);
        VFS_CHECK_LAUNCH();
    }
    long long ns = shell_count(d);
    if (d.ip || d.jp || d.kp) {
        hipLaunchKernelGGL(k_periodic_cellmetrics, dim3(grid_for(ns, 256)), dim3(256), 0, c->stream, d, c->g);
        VFS_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_face_metrics, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g);
    VFS_CHECK_LAUNCH();
    if (d.ip || d.jp || d.kp) {
        hipLaunchKernelGGL(k_periodic_allmetrics, dim3(grid_for(ns, 256)), dim3(256), 0, c->stream, d, c->g);
        VFS_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_aj_check, dim3(VFS_RED_BLOCKS), dim3(256), 0, c->stream, d, c->g.aj, c->red);
    VFS_CHECK_LAUNCH();
    double *res = c->red + (long long)VFS_RED_SLOTS * VFS_RED_BLOCKS;
    if (vfs_reduce_finish_slots(c, 0, VFS_RED_BLOCKS, 1, 0, res)) return -1;
    if (vfs_reduce_finish_slots(c, 1, VFS_RED_BLOCKS, 1, 1, res + 1)) return -1;
    double h[2];
    if (vfs_read_scalars(c, res, h, 2)) return -1;
    if (h[0] != 0.0) {
        const long long t = (long long)h[1];
        const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
        vfs_set_error("vfs_build_metrics: " + std::to_string((long long)h[0]) +
                      " cell(s) with a non-positive or non-finite Jacobian (a folded or degenerate grid), e.g. "
                      "i,j,k = " + std::to_string(i) + "," + std::to_string(j) + "," + std::to_string(k));
        return -7;
    }
    return 0;
}

extern "C" int vfs_get_metric(vfs_ctx *c, int which, double *dst)
{
    VFS_REQUIRE(which >= 0 && which <= 16, "metric index out of range");
    double *src[17] = {c->g.csi, c->g.eta, c->g.zet, c->g.aj, c->g.icsi, c->g.ieta, c->g.izet, c->g.iaj,
                       c->g.jcsi, c->g.jeta, c->g.jzet, c->g.jaj, c->g.kcsi, c->g.keta, c->g.kzet, c->g.kaj, c->g.cent};
    bool vec = !(which == 3 || which == 7 || which == 11 || which == 15);
    size_t n = (size_t)c->d.P * (vec ? 3 : 1);
    VFS_CHECK_HIP(hipMemcpyAsync(dst, src[which], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

// count of nvert values that FP32 does not hold exactly (per-block partials)
__global__ void k_nvert_f32_check(long long n, const double *__restrict__ nvert, double *part)
{
    __shared__ double sh[4];
    double acc = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        acc += (double)(float)nvert[t] != nvert[t] ? 1.0 : 0.0;
    double r = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

extern "C" int vfs_set_nvert(vfs_ctx *c, const double *nvert)
{
    c->nvert_version++;
    VFS_CHECK_HIP(hipMemcpyAsync(c->nvert, nvert, (size_t)c->d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (vfs_wall_mask(c)) return -1;   // the wall-function cells stay IB nodes
    hipLaunchKernelGGL(k_nvert_f32_check, dim3(VFS_RED_BLOCKS), dim3(256), 0, c->stream, c->d.P, c->nvert, c->red);
    VFS_CHECK_LAUNCH();
    double *res = c->red + (long long)VFS_RED_SLOTS * VFS_RED_BLOCKS;
    if (vfs_reduce_finish(c, VFS_RED_BLOCKS, 1, 0, res)) return -1;
    double bad = 0;
    if (vfs_read_scalars(c, res, &bad, 1)) return -1;
    c->nvert_f32 = bad == 0.0;
    return 0;
}

extern "C" int vfs_get_nvert(vfs_ctx *c, double *dst)
{
    VFS_CHECK_HIP(hipMemcpyAsync(dst, c->nvert, (size_t)c->d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// IB mask of an analytic sphere
// ---------------------------------------------------------------------------
// inside test at cell centres; 4 marks "inside this body" (ibm.c:224)
__global__ void k_sphere_inside(Dims d, const double *__restrict__ cent, double cx, double cy, double cz,
                                double r, double *nvert)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double v = 0.;
    if (i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2) {
        double dx = cent[3 * t] - cx, dy = cent[3 * t + 1] - cy, dz = cent[3 * t + 2] - cz;
        double dist = sqrt(dx * dx + dy * dy + dz * dz) - r;
        if (dist < 0.) v = 4.;
    }
    nvert[t] = v;
}

// near-boundary dilation over the clamped 3x3x3 box (ibm.c:319-341).  Reads
// only test for 4, which this pass never writes, so one parallel pass gives
// the sequential loop's result.
__global__ void k_sphere_dilate(Dims d, double *nvert)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double v = nvert[t];
    if (v < 0) v = 0;
    int ip = (i < d.mx - 1 ? (i + 1) : (i)), im = (i > 0 ? (i - 1) : (i));
    int jp = (j < d.my - 1 ? (j + 1) : (j)), jm = (j > 0 ? (j - 1) : (j));
    int kp = (k < d.mz - 1 ? (k + 1) : (k)), km = (k > 0 ? (k - 1) : (k));
    if ((int)(v + 0.5) != 4) {
        for (int kk = km; kk < kp + 1; kk++)
            for (int jj = jm; jj < jp + 1; jj++)
                for (int ii = im; ii < ip + 1; ii++)
                    if ((int)(nvert[IX(d, kk, jj, ii)] + 0.5) == 4) v = v > 2 ? v : 2;
        nvert[t] = v;
    }
}

// back to 3 (solid) / 1 (IB node) on interior cells (ibm.c:508-516)
__global__ void k_sphere_relabel(Dims d, double *nvert)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    double v = nvert[t];
    if ((int)(v + 0.5) == 2) v = 1;
    if ((int)(v + 0.5) == 4) v = 3;
    nvert[t] = v;
}

// the dilation and the relabelling that follow the inside test of any body
int vfs_mask_dilate_relabel(vfs_ctx *c)
{
    c->nvert_version++;
    const Dims d = c->d;
    unsigned gb = grid_for(d.P, 256);
    hipLaunchKernelGGL(k_sphere_dilate, dim3(gb), dim3(256), 0, c->stream, d, c->nvert);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sphere_relabel, dim3(gb), dim3(256), 0, c->stream, d, c->nvert);
    VFS_CHECK_LAUNCH();
    return 0;
}

extern "C" int vfs_nvert_sphere(vfs_ctx *c, double cx, double cy, double cz, double r, double *nvert_out)
{
    c->nvert_version++;
    const Dims d = c->d;
    unsigned gb = grid_for(d.P, 256);
    hipLaunchKernelGGL(k_sphere_inside, dim3(gb), dim3(256), 0, c->stream, d, c->g.cent, cx, cy, cz, r, c->nvert);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sphere_dilate, dim3(gb), dim3(256), 0, c->stream, d, c->nvert);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sphere_relabel, dim3(gb), dim3(256), 0, c->stream, d, c->nvert);
    VFS_CHECK_LAUNCH();
    c->nvert_f32 = 1;   // values 0, 1, 3
    if (nvert_out)
        VFS_CHECK_HIP(hipMemcpyAsync(nvert_out, c->nvert, (size_t)d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}
