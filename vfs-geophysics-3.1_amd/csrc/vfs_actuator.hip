// Actuator models around the hot path: the element search of the rotor
// pre-processing (rotor_model.c:2590-2737) and the spreading of Lagrangian
// element forces onto the grid, Calc_F_eul (rotor_model.c:4792-5137), with
// its discrete delta functions (rotor_model.c:6042-6200).  The spread force
// F_eul is added to the implicit residual last (momentum.c:2288) once it is
// attached with vfs_set_body_force.
//
// Elements are arrays of n on the device, in blocks: centres x|y|z,
// forces x|y|z, areas, index boxes i_min|i_max|j_min|j_max|k_min|k_max
// (upper bounds exclusive), widths x|y|z.
//
// The spreading gathers instead of scattering: one thread a cell of the
// boxes' union, adding the elements whose box holds the cell in element
// order, which is the reference's order of accumulation (no atomics).
#include <vector>

#include "vfs_internal.h"

// rotor_model.c:6085-6197
__device__ __forceinline__ double dfunc_dev(int df, double r, double n)
{
    const double pi = 3.14159265358979323846;   // M_PI (dfunc_sc4h)
    const double pi8 = 3.14159265;              // the truncated pi of dfunc_s4h
    const double ar = fabs(r);
    switch (df) {
    case 0: return ar < 1.0 ? 1.0 - ar : 0.0;
    case 6: return ar < n ? (n - ar) / (n * n) : 0.0;
    case 7: return exp(-(r / n) * (r / n)) / (n * sqrt(3.1415926));
    case 8:
        if (ar <= 0.5) return 3.0 / 8.0 + pi8 / 32.0 - r * r / 4.0;
        if (ar >= 0.5 && ar <= 1.5)
            return 1.0 / 4.0 + (1.0 - ar) * sqrt(-2.0 + 8.0 * ar - 4.0 * (r * r)) / 8.0 -
                   asin(sqrt(2.0) * (ar - 1.0)) / 8.0;
        if (ar >= 1.5 && ar <= 2.5)
            return 17.0 / 16.0 - pi8 / 64.0 - 3.0 * ar / 4.0 + r * r / 8.0 +
                   (ar - 2.0) * sqrt(-14.0 + 16.0 * ar - 4.0 * (r * r)) / 16.0 + asin(sqrt(2.0) * (ar - 2.0)) / 16.0;
        return 0.0;
    default:
        if (ar <= 1.5) return 0.25 * (pi + 2 * sin(0.25 * pi * (2 * ar + 1)) - 2 * sin(0.25 * pi * (2 * ar - 1))) / pi;
        if (ar >= 1.5 && ar <= 2.5) return -0.125 * (-5 * pi + 2 * pi * ar + 4 * sin(0.25 * pi * (2 * ar - 1))) / pi;
        return 0.0;
    }
}

// first minimum of (d, order) over a wavefront: the smaller distance, and on
// a tie the earlier loop position
__device__ __forceinline__ void wave_first_min(double &d, long long &o)
{
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        const double d2 = __shfl_xor(d, s);
        const long long o2 = __shfl_xor(o, s);
        if (d2 < d || (d2 == d && o2 < o)) { d = d2; o = o2; }
    }
}

// first minimum of (d, order) over a workgroup of RE_WAVES wavefronts
constexpr int RE_WAVES = 4;
__device__ __forceinline__ void block_first_min(double &d, long long &o, double *sd, long long *so)
{
    wave_first_min(d, o);
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) sd[w] = d, so[w] = o;
    __syncthreads();
    d = sd[0], o = so[0];
    for (int q = 1; q < RE_WAVES; q++)
        if (sd[q] < d || (sd[q] == d && so[q] < o)) d = sd[q], o = so[q];
    __syncthreads();
}

// One workgroup per element: the closest node, first among every 20th node
// in i, j, k loop order, then in the +-21 window around it; the widths of
// that cell and the box +- itpwidth clamped to the interior.  The first
// minimum (smallest distance, then earliest loop position) does not depend
// on how the positions are spread over the threads.
__global__ void __launch_bounds__(64 * RE_WAVES) k_rotor_elements(Dims d, Geom g, const double *__restrict__ coor,
                                                                  int n, const double *__restrict__ cent,
                                                                  int itpwidth, int *__restrict__ box,
                                                                  double *__restrict__ dh)
{
    __shared__ double sd[RE_WAVES];
    __shared__ long long so[RE_WAVES];
    const int l = blockIdx.x, lane = threadIdx.x;
    if (l >= n) return;
    const int lxs = 1, lxe = d.mx - 1, lys = 1, lye = d.my - 1, lzs = 1, lze = d.mz - 1;
    const double cx = cent[l], cy = cent[n + l], cz = cent[2 * n + l];
    auto dist = [&](int i, int j, int k) {
        const long long id = IX(d, k, j, i);
        const double r1 = cx - coor[3 * id], r2 = cy - coor[3 * id + 1], r3 = cz - coor[3 * id + 2];
        return sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    };
    // coarse: i = lxs + 20 a, j = lys + 20 b, k = lzs + 20 c, order (a, b, c)
    const int na = (lxe - lxs + 19) / 20, nb = (lye - lys + 19) / 20, nc = (lze - lzs + 19) / 20;
    double dm = 1.e6;
    long long om = -1;
    for (long long o = lane; o < (long long)na * nb * nc; o += 64 * RE_WAVES) {
        const int a = (int)(o / ((long long)nb * nc)), b = (int)((o / nc) % nb), c = (int)(o % nc);
        const double d1 = dist(lxs + 20 * a, lys + 20 * b, lzs + 20 * c);
        if (d1 < dm) { dm = d1; om = o; }
    }
    block_first_min(dm, om, sd, so);
    int icc = lxs, jcc = lys, kcc = lzs;
    if (om >= 0) {
        icc = lxs + 20 * (int)(om / ((long long)nb * nc));
        jcc = lys + 20 * (int)((om / nc) % nb);
        kcc = lzs + 20 * (int)(om % nc);
    }
    const int i0 = max(icc - 21, lxs), i1 = min(icc + 21, lxe), j0 = max(jcc - 21, lys), j1 = min(jcc + 21, lye);
    const int k0 = max(kcc - 21, lzs), k1 = min(kcc + 21, lze);
    const int wi = max(i1 - i0, 0), wj = max(j1 - j0, 0), wk = max(k1 - k0, 0);
    dm = 1.e13;
    om = -1;
    for (long long o = lane; o < (long long)wi * wj * wk; o += 64 * RE_WAVES) {
        const int a = (int)(o / ((long long)wj * wk)), b = (int)((o / wk) % wj), c = (int)(o % wk);
        const double d1 = dist(i0 + a, j0 + b, k0 + c);
        if (d1 < dm) { dm = d1; om = o; }
    }
    block_first_min(dm, om, sd, so);
    if (lane != 0) return;
    int ic = 0, jc = 0, kc = 0;   // the reference's initial values when the window is empty
    if (om >= 0) {
        ic = i0 + (int)(om / ((long long)wj * wk));
        jc = j0 + (int)((om / wk) % wj);
        kc = k0 + (int)(om % wk);
    }
    const long long id = IX(d, kc, jc, ic);
    const double *a = g.csi + 3 * id, *b = g.eta + 3 * id, *e = g.zet + 3 * id;
    double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    dh[l] = 1.0 / g.aj[id] / area;
    area = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    dh[n + l] = 1.0 / g.aj[id] / area;
    area = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    dh[2 * n + l] = 1.0 / g.aj[id] / area;
    const int cc[3] = {ic, jc, kc}, lo[3] = {lxs, lys, lzs}, hi[3] = {lxe, lye, lze};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int c1 = cc[q] - itpwidth, c2 = cc[q] + itpwidth;
        if (c1 > hi[q] || c2 < lo[q]) {
            box[2 * q * n + l] = lo[q];
            box[(2 * q + 1) * n + l] = lo[q];
        } else {
            box[2 * q * n + l] = max(c1, lo[q]);
            box[(2 * q + 1) * n + l] = min(c2, hi[q]);
        }
    }
}

extern "C" int vfs_rotor_elements(vfs_ctx *c, const double *coor, int n, const double *cent, int itpwidth, int *box,
                                  double *dh)
{
    VFS_REQUIRE(n >= 0 && (n == 0 || (coor && cent && box && dh)), "bad element arrays");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_rotor_elements, dim3(n), dim3(64 * RE_WAVES), 0, c->stream, c->d, c->g, coor, n, cent, itpwidth, box, dh);
    VFS_CHECK_LAUNCH();
    return 0;
}

// union of the element boxes: lo/hi of i, j, k (empty boxes are skipped)
__global__ void k_box_union(int n, const int *__restrict__ box, int *__restrict__ u)
{
    __shared__ int s[6][256];
    int v[6] = {1 << 30, -1, 1 << 30, -1, 1 << 30, -1};
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
        const int b[6] = {box[l], box[n + l], box[2 * n + l], box[3 * n + l], box[4 * n + l], box[5 * n + l]};
        if (b[0] >= b[1] || b[2] >= b[3] || b[4] >= b[5]) continue;
        for (int q = 0; q < 3; q++) {
            v[2 * q] = min(v[2 * q], b[2 * q]);
            v[2 * q + 1] = max(v[2 * q + 1], b[2 * q + 1]);
        }
    }
    for (int q = 0; q < 6; q++) s[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 6; q++)
                s[q][threadIdx.x] = (q & 1) ? max(s[q][threadIdx.x], s[q][threadIdx.x + w])
                                            : min(s[q][threadIdx.x], s[q][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) u[threadIdx.x] = s[threadIdx.x][0];
}

// rotor_model.c:2550-2583 Calculate_normal: unit covariant base vectors
__device__ __forceinline__ void normals(const double *csi, const double *eta, const double *zet, double ni[3],
                                        double nj[3], double nk[3])
{
    const double a11 = csi[0], a12 = csi[1], a13 = csi[2];
    const double a21 = eta[0], a22 = eta[1], a23 = eta[2];
    const double a31 = zet[0], a32 = zet[1], a33 = zet[2];
    const double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    double x[3] = {(a33 * a22 - a32 * a23) / det, -(a33 * a21 - a31 * a23) / det, (a32 * a21 - a31 * a22) / det};
    double y[3] = {-(a33 * a12 - a32 * a13) / det, (a33 * a11 - a31 * a13) / det, -(a32 * a11 - a31 * a12) / det};
    double z[3] = {(a23 * a12 - a22 * a13) / det, -(a23 * a11 - a21 * a13) / det, (a22 * a11 - a21 * a12) / det};
    const double si = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double sj = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    const double sk = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
#pragma unroll
    for (int q = 0; q < 3; q++) { ni[q] = x[q] / si; nj[q] = y[q] / sj; nk[q] = z[q] / sk; }
}

// rotor_model.c:4879-4994 for the cells of the box union: one thread a
// cell, elements in order.  A workgroup first lists, in order, the elements
// of each batch of 256 whose box meets the box of its cells; a cell then
// runs the reference's per-element test on those only, so its sum takes the
// same terms in the same order.
constexpr int FEUL_BLOCK = 256;
__global__ void __launch_bounds__(FEUL_BLOCK) k_f_eul(Dims d, Geom g, const double *__restrict__ coor, int n,
                                                      const double *__restrict__ cent, const double *__restrict__ F,
                                                      const double *__restrict__ dA, const int *__restrict__ box,
                                                      int df, double halfwidth, int width_mode, double fx, double fy,
                                                      double fz, int i0, int j0, int k0, int ni_, int nj_, int nk_,
                                                      double *__restrict__ f_eul)
{
    __shared__ int sl[FEUL_BLOCK];
    const long long nt = (long long)ni_ * nj_ * nk_;
    const long long t0 = (long long)blockIdx.x * FEUL_BLOCK, t = t0 + threadIdx.x;
    // the box of the workgroup's cells (whole rows / planes once it wraps)
    int bi0, bi1, bj0, bj1, bk0, bk1;
    {
        const long long tl = (t0 + FEUL_BLOCK < nt ? t0 + FEUL_BLOCK : nt) - 1;
        const int a0 = (int)(t0 % ni_), b0 = (int)((t0 / ni_) % nj_), c0 = (int)(t0 / ((long long)ni_ * nj_));
        const int a1 = (int)(tl % ni_), b1 = (int)((tl / ni_) % nj_), c1 = (int)(tl / ((long long)ni_ * nj_));
        bk0 = k0 + c0, bk1 = k0 + c1;
        if (c0 == c1) bj0 = j0 + b0, bj1 = j0 + b1;
        else bj0 = j0, bj1 = j0 + nj_ - 1;
        if (c0 == c1 && b0 == b1) bi0 = i0 + a0, bi1 = i0 + a1;
        else bi0 = i0, bi1 = i0 + ni_ - 1;
    }
    const bool live = t < nt;
    const long long tc = live ? t : t0;   // threads past the last cell work on the first one and write nothing
    const int i = i0 + (int)(tc % ni_), j = j0 + (int)((tc / ni_) % nj_), k = k0 + (int)(tc / ((long long)ni_ * nj_));
    const long long id = IX(d, k, j, i);
#define CO(kk, jj, ii, q) coor[3 * IX(d, kk, jj, ii) + (q)]
    double xi[3], xj[3], xk[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        xi[q] = (CO(k, j, i, q) + CO(k - 1, j, i, q) + CO(k, j - 1, i, q) + CO(k - 1, j - 1, i, q)) * 0.25;
        xj[q] = (CO(k, j, i, q) + CO(k - 1, j, i, q) + CO(k, j, i - 1, q) + CO(k - 1, j, i - 1, q)) * 0.25;
        xk[q] = (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k, j, i - 1, q) + CO(k, j - 1, i - 1, q)) * 0.25;
    }
#undef CO
    const double *cs = g.csi + 3 * id, *et = g.eta + 3 * id, *ze = g.zet + 3 * id;
    double ni[3], nj[3], nk[3];
    normals(cs, et, ze, ni, nj, nk);
    double dhm[3];
    {
        double area = sqrt(cs[0] * cs[0] + cs[1] * cs[1] + cs[2] * cs[2]);
        dhm[0] = 1.0 / g.aj[id] / area;
        area = sqrt(et[0] * et[0] + et[1] * et[1] + et[2] * et[2]);
        dhm[1] = 1.0 / g.aj[id] / area;
        area = sqrt(ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
        dhm[2] = 1.0 / g.aj[id] / area;
    }
    const double *m[3] = {cs, et, ze};
    double acc[3] = {0.0, 0.0, 0.0};
    for (int l0 = 0; l0 < n; l0 += FEUL_BLOCK) {
        // order-keeping compaction of the batch's elements that meet the box
        const int le = l0 + (int)threadIdx.x;
        const int meet = le < n && !(bi1 < box[le] || bi0 >= box[n + le] || bj1 < box[2 * n + le] ||
                                     bj0 >= box[3 * n + le] || bk1 < box[4 * n + le] || bk0 >= box[5 * n + le]);
        sl[threadIdx.x] = meet;
        __syncthreads();
        for (int w = 1; w < FEUL_BLOCK; w <<= 1) {   // inclusive scan
            const int v = threadIdx.x >= (unsigned)w ? sl[threadIdx.x - w] : 0;
            __syncthreads();
            sl[threadIdx.x] += v;
            __syncthreads();
        }
        const int cnt = sl[FEUL_BLOCK - 1], pos = sl[threadIdx.x] - 1;
        __syncthreads();
        if (meet) sl[pos] = le;
        __syncthreads();
        for (int ql = 0; live && ql < cnt; ql++) {
        const int l = sl[ql];
        if (i < box[l] || i >= box[n + l] || j < box[2 * n + l] || j >= box[3 * n + l] || k < box[4 * n + l] ||
            k >= box[5 * n + l])
            continue;
        const double cx = cent[l], cy = cent[n + l], cz = cent[2 * n + l], a = dA[l];
        double dhx, dhy, dhz;
        if (width_mode == 1) {
            dhx = fx; dhy = fy; dhz = fz;
        } else if (width_mode == 2 || width_mode == 3) {
            dhx = width_mode == 3 ? sqrt(a) : a;
            dhy = dhx; dhz = dhx;
        } else {
            dhx = dhm[0]; dhy = dhm[1]; dhz = dhm[2];
        }
        const double rx[3] = {xi[0] - cx, xj[0] - cx, xk[0] - cx};
        const double ry[3] = {xi[1] - cy, xj[1] - cy, xk[1] - cy};
        const double rz[3] = {xi[2] - cz, xj[2] - cz, xk[2] - cz};
        const double vol_eul = 1.0 / (dhx * dhy * dhz);
        const double Fx = F[l], Fy = F[n + l], Fz = F[2 * n + l];
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const double r1 = fabs(rx[f] * ni[0] + ry[f] * ni[1] + rz[f] * ni[2]) / dhx;
            const double r2 = fabs(rx[f] * nj[0] + ry[f] * nj[1] + rz[f] * nj[2]) / dhy;
            const double r3 = fabs(rx[f] * nk[0] + ry[f] * nk[1] + rz[f] * nk[2]) / dhz;
            const double dfn = vol_eul * dfunc_dev(df, r1, halfwidth) * dfunc_dev(df, r2, halfwidth) *
                               dfunc_dev(df, r3, halfwidth);
            acc[f] += Fx * dfn * a * m[f][0] + Fy * dfn * a * m[f][1] + Fz * dfn * a * m[f][2];
        }
        }
        __syncthreads();   // sl is rewritten by the next batch
    }
    if (!live) return;
    f_eul[3 * id] = acc[0];
    f_eul[3 * id + 1] = acc[1];
    f_eul[3 * id + 2] = acc[2];
}

// rotor_model.c:5006-5118 on every point: zero force where the 3x3x3 nvert
// sum (clamped to the interior) exceeds 2.9; then the folds of the ghost
// values into the first and last cells, which add zeros (no element box
// reaches a ghost point)
__global__ void k_f_eul_finish(Dims d, const double *__restrict__ nvert, double *__restrict__ f_eul)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double f[3] = {f_eul[3 * t], f_eul[3 * t + 1], f_eul[3 * t + 2]};
    if (i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2) {
        double nv = 0.0;
        for (int kk = max(k - 1, 1); kk < min(k + 2, d.mz - 1); kk++)
            for (int jj = max(j - 1, 1); jj < min(j + 2, d.my - 1); jj++)
                for (int ii = max(i - 1, 1); ii < min(i + 2, d.mx - 1); ii++) nv += nvert[IX(d, kk, jj, ii)];
        if (nv > 2.9) f[0] = f[1] = f[2] = 0.0;
    }
    const int folds = (i == 1) + (i == d.mx - 2) + (j == 1) + (j == d.my - 2) + (k == 1) + (k == d.mz - 2);
    for (int s = 0; s < folds; s++) { f[0] += 0.0; f[1] += 0.0; f[2] += 0.0; }
    f_eul[3 * t] = f[0];
    f_eul[3 * t + 1] = f[1];
    f_eul[3 * t + 2] = f[2];
}

__global__ void k_zero(long long n, double *__restrict__ x)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = 0.0;
}

extern "C" int vfs_calc_f_eul(vfs_ctx *c, const double *coor, int n, const double *cent, const double *F,
                              const double *dA, const int *box, int df, double halfwidth, int width_mode,
                              const double *dh_fixed, double *f_eul)
{
    const Dims d = c->d;
    VFS_REQUIRE(n >= 0 && f_eul && (n == 0 || (coor && cent && F && dA && box)), "bad element arrays");
    VFS_REQUIRE(width_mode >= 0 && width_mode <= 3, "width_mode must be 0 (metrics), 1 (fixed), 2 (dA) or 3 (sqrt dA)");
    VFS_REQUIRE(width_mode != 1 || dh_fixed, "width_mode 1 needs dh_fixed (3 doubles, host)");
    VFS_REQUIRE(df != 6 && df != 7 ? true : halfwidth > 0, "delta functions 6 and 7 need halfwidth > 0");
    hipLaunchKernelGGL(k_zero, dim3(grid_for(3 * d.P, 256)), dim3(256), 0, c->stream, 3 * d.P, f_eul);
    VFS_CHECK_LAUNCH();
    if (n > 0) {
        int *u = nullptr;
        VFS_CHECK_HIP(hipMalloc(&u, 6 * sizeof(int)));
        hipLaunchKernelGGL(k_box_union, dim3(1), dim3(256), 0, c->stream, n, box, u);
        int h[6];
        const hipError_t e1 = hipGetLastError();
        const hipError_t e2 = e1 == hipSuccess ? hipMemcpyAsync(h, u, sizeof h, hipMemcpyDeviceToHost, c->stream) : e1;
        const hipError_t e3 = e2 == hipSuccess ? hipStreamSynchronize(c->stream) : e2;
        hipFree(u);
        VFS_CHECK_HIP(e3);
        if (h[1] >= 0) {   // some box is not empty
            VFS_REQUIRE(h[0] >= 1 && h[2] >= 1 && h[4] >= 1 && h[1] <= d.mx - 1 && h[3] <= d.my - 1 && h[5] <= d.mz - 1,
                        "element boxes must lie in the interior (as vfs_rotor_elements makes them)");
            const int ni = h[1] - h[0], nj = h[3] - h[2], nk = h[5] - h[4];
            const long long cells = (long long)ni * nj * nk;
            const double fx = width_mode == 1 ? dh_fixed[0] : 0, fy = width_mode == 1 ? dh_fixed[1] : 0,
                         fz = width_mode == 1 ? dh_fixed[2] : 0;
            hipLaunchKernelGGL(k_f_eul, dim3(grid_for(cells, FEUL_BLOCK)), dim3(FEUL_BLOCK), 0, c->stream, d, c->g, coor, n, cent,
                               F, dA, box, df, halfwidth, width_mode, fx, fy, fz, h[0], h[2], h[4], ni, nj, nk, f_eul);
            VFS_CHECK_LAUNCH();
        }
    }
    hipLaunchKernelGGL(k_f_eul_finish, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->nvert, f_eul);
    VFS_CHECK_LAUNCH();
    return 0;
}

extern "C" int vfs_set_body_force(vfs_ctx *c, const double *f_eul)
{
    c->body_force = f_eul;
    return 0;
}
