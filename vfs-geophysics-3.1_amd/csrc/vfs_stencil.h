// Per-point stencil bodies of the momentum right-hand side, shared by the
// reference-structured passes and the fused kernel.  Each function keeps the
// reference's expression and operation order (cited per function).
#pragma once
#include "vfs_internal.h"

#define VFS_SOLID 0.1 // k-omega.c:11

// face-centred velocity gradients (k-omega.c:48-278).
// The face joins cell c0 = (k,j,i) and c1 = c0 + e_dir.  The normal
// derivative is U(c1) - U(c0).  Each transverse derivative (direction t) is
//   (U(c1 + p e_t) + U(c0 + p e_t) - U(c1 + m e_t) - U(c0 + m e_t)) * s
// with (p, m, s) = (0, -1, 0.5) when a cell on the + side is solid,
// (1, 0, 0.5) when one on the - side is, and (1, -1, 0.25) otherwise
// (k-omega.c:74-104 and the analogous j/k branches).  Selecting (p, m, s)
// first and reading once gives the same expression as each branch of the
// reference, in the same order.
template <typename S, typename Fetch, typename NFetch>
__device__ __forceinline__ void du_face(int dir, const Fetch &U, const NFetch &N, const Dims &d,
                                        int i, int j, int k, S dc[3], S de[3], S dz[3])
{
    (void)d;
    // unit offsets of the face normal and the two transverse directions
    const int ni = dir == 0, nj = dir == 1, nk = dir == 2;
    const int i1 = i + ni, j1 = j + nj, k1 = k + nk;   // c1
    S *dn = dir == 0 ? dc : dir == 1 ? de : dz;
#pragma unroll
    for (int q = 0; q < 3; q++) dn[q] = U(k1, j1, i1, q) - U(k, j, i, q);
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
        // transverse direction: the two of {i, j, k} other than dir, in order
        const int t = dir == 0 ? 1 + tt : dir == 1 ? 2 * tt : tt;
        const int ti = t == 0, tj = t == 1, tk = t == 2;
        int p = 1, m = -1;
        double sc = 0.25;
        if (N(k + tk, j + tj, i + ti) > VFS_SOLID || N(k1 + tk, j1 + tj, i1 + ti) > VFS_SOLID) {
            p = 0; m = -1; sc = 0.5;
        } else if (N(k - tk, j - tj, i - ti) > VFS_SOLID || N(k1 - tk, j1 - tj, i1 - ti) > VFS_SOLID) {
            p = 1; m = 0; sc = 0.5;
        }
        S *dt = t == 0 ? dc : t == 1 ? de : dz;
#pragma unroll
        for (int q = 0; q < 3; q++)
            dt[q] = (U(k1 + p * tk, j1 + p * tj, i1 + p * ti, q) + U(k + p * tk, j + p * tj, i + p * ti, q) -
                     U(k1 + m * tk, j1 + m * tj, i1 + m * ti, q) - U(k + m * tk, j + m * tj, i + m * ti, q)) *
                    sc;
    }
}

// cell-centred gradient for the LES model (k-omega.c:305-423, k==0 quirk at 391)
__device__ __forceinline__ void du_center(const Dims &d, int i, int j, int k, const double *__restrict__ ucat,
                                          const double *__restrict__ nvert, double dc[3], double de[3], double dz[3])
{
#define U3(kk, jj, ii, q) ucat[3 * IX(d, kk, jj, ii) + (q)]
#define NV(kk, jj, ii) nvert[IX(d, kk, jj, ii)]
    if (NV(k, j, i + 1) > VFS_SOLID || (!d.ip && i == d.mx - 2)) {
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j, i, q) - U3(k, j, i - 1, q));
    } else if (NV(k, j, i - 1) > VFS_SOLID || (!d.ip && i == 1)) {
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j, i + 1, q) - U3(k, j, i, q));
    } else {
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j, i + 1, q) - U3(k, j, i - 1, q)) * 0.5;
    }
    if (NV(k, j + 1, i) > VFS_SOLID || (!d.jp && j == d.my - 2)) {
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j, i, q) - U3(k, j - 1, i, q));
    } else if (NV(k, j - 1, i) > VFS_SOLID || (!d.jp && j == 1)) {
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j + 1, i, q) - U3(k, j, i, q));
    } else {
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j + 1, i, q) - U3(k, j - 1, i, q)) * 0.5;
    }
    if (NV(k + 1, j, i) > VFS_SOLID || (!d.kp && k == d.mz - 2)) {
        for (int q = 0; q < 3; q++) dz[q] = (U3(k, j, i, q) - U3(k - 1, j, i, q));
    } else if (NV(k - 1, j, i) > VFS_SOLID || (!d.kp && k == 0)) {
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j, i, q) - U3(k, j, i, q));
    } else {
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j, i, q) - U3(k - 1, j, i, q)) * 0.5;
    }
#undef U3
#undef NV
}

// Dual number (value, tangent) for the tangent-linear (JVP) instantiation of
// the face fluxes; |x| differentiates to sign(x) (0 at x = 0).
struct Dual {
    double v, d;
    __device__ __forceinline__ Dual() : v(0), d(0) {}
    __device__ __forceinline__ Dual(double a) : v(a), d(0) {}
    __device__ __forceinline__ Dual(double a, double b) : v(a), d(b) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual fabs(Dual a) { return Dual(fabs(a.v), a.v > 0 ? a.d : (a.v < 0 ? -a.d : 0.0)); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
__device__ __forceinline__ Dual operator/(double a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }
__device__ __forceinline__ double value_of(double a) { return a; }
__device__ __forceinline__ double value_of(Dual a) { return a.v; }

// level.c:1734-1771 weno3; the reference's pow(x, 2.) is x * x (its -O3
// build folds it so).  The upwind side follows the sign of the wave speed.
template <typename S>
__device__ __forceinline__ S weno3(S f0, S f1, S f2, S f3, S wavespeed)
{
    S fL, fC, fR;
    if (value_of(wavespeed) > 0) {
        fL = f0; fC = f1; fR = f2;
    } else {
        fL = f3; fC = f2; fR = f1;
    }
    const double d0 = 2. / 3., d1 = 1. / 3.;
    const double eps = 1.e-6;
    const S b0 = fC - fR, b1 = fL - fC;
    const S beta0 = b0 * b0, beta1 = b1 * b1;
    const S e0 = eps + beta0, e1 = eps + beta1;
    const S alpha0 = d0 / (e0 * e0), alpha1 = d1 / (e1 * e1);
    const S sumalpha = alpha0 + alpha1;
    const S w0 = alpha0 / sumalpha, w1 = alpha1 / sumalpha;
    const S u0 = (fC * 0.5 + fR * 0.5);
    const S u1 = (-fL * 0.5 + fC * 1.5);
    return w0 * u0 + w1 * u1;
}

struct GlobalSca {
    const double *__restrict__ p;
    Dims d;
    __device__ __forceinline__ double operator()(int k, int j, int i) const { return p[IX(d, k, j, i)]; }
};

struct GlobalVec {
    const double *__restrict__ p;
    Dims d;
    __device__ __forceinline__ double operator()(int k, int j, int i, int q) const { return p[3 * IX(d, k, j, i) + q]; }
};
struct DualVec {
    const double *__restrict__ p, *__restrict__ t;
    Dims d;
    __device__ __forceinline__ Dual operator()(int k, int j, int i, int q) const
    {
        long long o = 3 * IX(d, k, j, i) + q;
        return Dual(p[o], t[o]);
    }
};

// Convective (div) and viscous (visc) flux through face `dir` of point
// (k,j,i): momentum.c:672-931 (i), 934-1166 (j), 1169-1409 (k).
// S = double for the residual, Dual for its tangent; U fetches ucat, UCon
// fetches the DIR component of ucont at a DA index.
// Face metrics of the DIR face at DA index id, read from the stored arrays.
struct StoredFaceMetric {
    template <int DIR, typename Idx>
    __device__ __forceinline__ void get(const Dims &, const Geom &g, Idx id, double cs[3], double et[3],
                                        double ze[3], double &aj) const
    {
        const double *fc = DIR == 0 ? g.icsi : DIR == 1 ? g.jcsi : g.kcsi;
        const double *fe = DIR == 0 ? g.ieta : DIR == 1 ? g.jeta : g.keta;
        const double *fz = DIR == 0 ? g.izet : DIR == 1 ? g.jzet : g.kzet;
        const double *faj = DIR == 0 ? g.iaj : DIR == 1 ? g.jaj : g.kaj;
        aj = ldx(faj, id);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            cs[q] = ldx(fc, (Idx)(3 * id + q)); et[q] = ldx(fe, (Idx)(3 * id + q)); ze[q] = ldx(fz, (Idx)(3 * id + q));
        }
    }
};

// A mask reader for faces with no marked cell in reach: every nvert test of
// the face is false, so the compiler removes them.
struct ZeroN {
    __device__ __forceinline__ double operator()(int, int, int) const { return 0.0; }
};

// Calculate_normal_and_area + Calculate_dxdydz (rhs2.c:592-643): the cell's
// length scales along x, y and z from the metrics cs / et / ze and the
// Jacobian ajc (volume 1 / ajc divided by the face areas, along the unit
// covariant base vectors).
__device__ __forceinline__ void dxdydz(double ajc, const double cs[3], const double et[3], const double ze[3],
                                       double &dx, double &dy, double &dz)
{
    const double vol = 1. / ajc;
    const double a11 = cs[0], a12 = cs[1], a13 = cs[2];
    const double a21 = et[0], a22 = et[1], a23 = et[2];
    const double a31 = ze[0], a32 = ze[1], a33 = ze[2];
    const double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    double nx_i = (a33 * a22 - a32 * a23) / det, nx_j = -(a33 * a12 - a32 * a13) / det, nx_k = (a23 * a12 - a22 * a13) / det;
    double ny_i = -(a33 * a21 - a31 * a23) / det, ny_j = (a33 * a11 - a31 * a13) / det, ny_k = -(a23 * a11 - a21 * a13) / det;
    double nz_i = (a32 * a21 - a31 * a22) / det, nz_j = -(a32 * a11 - a31 * a12) / det, nz_k = (a22 * a11 - a21 * a12) / det;
    const double sum_i = sqrt(nx_i * nx_i + ny_i * ny_i + nz_i * nz_i);
    const double sum_j = sqrt(nx_j * nx_j + ny_j * ny_j + nz_j * nz_j);
    const double sum_k = sqrt(nx_k * nx_k + ny_k * ny_k + nz_k * nz_k);
    const double Ai = sqrt(a11 * a11 + a12 * a12 + a13 * a13);
    const double Aj = sqrt(a21 * a21 + a22 * a22 + a23 * a23);
    const double Ak = sqrt(a31 * a31 + a32 * a32 + a33 * a33);
    nx_i /= sum_i, ny_i /= sum_i, nz_i /= sum_i;
    nx_j /= sum_j, ny_j /= sum_j, nz_j /= sum_j;
    nx_k /= sum_k, ny_k /= sum_k, nz_k /= sum_k;
    const double Li = vol / Ai, Lj = vol / Aj, Lk = vol / Ak;
    dx = fabs(Li * nx_i + Lj * nx_j + Lk * nx_k);
    dy = fabs(Li * ny_i + Lj * ny_j + Lk * ny_k);
    dz = fabs(Li * nz_i + Lj * nz_j + Lk * nz_k);
}

// WENO: false compiles the WENO3 branch out (for kernels launched only when
// the weno flag is off).
// Idx: index type of the global arrays (unsigned: 32-bit offsets).
// PLAIN: the caller guarantees that no cell the face reads is marked
// (nvert = 0) and that the face is neither the first nor the last of its
// direction (0 < n < m-2); every branch those conditions decide is then the
// interior one, and the checks are compiled out.  Same expressions, so the
// results are bit-identical to the general path on such faces.
// OPTS: the -skew and -clark terms are compiled in (A: the advective flux of
// -skew, written when d.skew is set).
template <int DIR, typename S, typename FU, typename FC, typename FN, typename FM = StoredFaceMetric, bool WENO = true,
          typename Idx = long long, bool PLAIN = false, bool OPTS = false>
__device__ __forceinline__ void face_flux_t(const Dims &d, const Geom &g, const FN &N,
                                            const FC &UCon, const FU &U, const double *__restrict__ nu_t,
                                            int i, int j, int k, S D[3], S V[3], const FM &fm = FM(), S *A = nullptr)
{
    const Idx id = ixt<Idx>(d, k, j, i);
    double mcs[3], met[3], mze[3], ajc;
    fm.template get<DIR>(d, g, id, mcs, met, mze, ajc);
    const double csi0 = mcs[0], csi1 = mcs[1], csi2 = mcs[2];
    const double eta0 = met[0], eta1 = met[1], eta2 = met[2];
    const double zet0 = mze[0], zet1 = mze[1], zet2 = mze[2];
    S dc[3], de[3], dz[3];
    du_face<S>(DIR, U, N, d, i, j, k, dc, de, dz);
    double g11, g21, g31, n0, n1, n2;
    if (DIR == 0) {
        g11 = csi0 * csi0 + csi1 * csi1 + csi2 * csi2;
        g21 = eta0 * csi0 + eta1 * csi1 + eta2 * csi2;
        g31 = zet0 * csi0 + zet1 * csi1 + zet2 * csi2;
        n0 = csi0, n1 = csi1, n2 = csi2;
    } else if (DIR == 1) {
        g11 = csi0 * eta0 + csi1 * eta1 + csi2 * eta2;
        g21 = eta0 * eta0 + eta1 * eta1 + eta2 * eta2;
        g31 = zet0 * eta0 + zet1 * eta1 + zet2 * eta2;
        n0 = eta0, n1 = eta1, n2 = eta2;
    } else {
        g11 = csi0 * zet0 + csi1 * zet1 + csi2 * zet2;
        g21 = eta0 * zet0 + eta1 * zet1 + eta2 * zet2;
        g31 = zet0 * zet0 + zet1 * zet1 + zet2 * zet2;
        n0 = zet0, n1 = zet1, n2 = zet2;
    }
    const S r11 = dc[0] * csi0 + de[0] * eta0 + dz[0] * zet0;
    const S r21 = dc[1] * csi0 + de[1] * eta0 + dz[1] * zet0;
    const S r31 = dc[2] * csi0 + de[2] * eta0 + dz[2] * zet0;
    const S r12 = dc[0] * csi1 + de[0] * eta1 + dz[0] * zet1;
    const S r22 = dc[1] * csi1 + de[1] * eta1 + dz[1] * zet1;
    const S r32 = dc[2] * csi1 + de[2] * eta1 + dz[2] * zet1;
    const S r13 = dc[0] * csi2 + de[0] * eta2 + dz[0] * zet2;
    const S r23 = dc[1] * csi2 + de[1] * eta2 + dz[1] * zet2;
    const S r33 = dc[2] * csi2 + de[2] * eta2 + dz[2] * zet2;

    const int n = DIR == 0 ? i : DIR == 1 ? j : k;
    const int m = DIR == 0 ? d.mx : DIR == 1 ? d.my : d.mz;
    const int per = DIR == 0 ? d.ip : DIR == 1 ? d.jp : d.kp;
    const Idx st = DIR == 0 ? (Idx)1 : DIR == 1 ? (Idx)d.mx : (Idx)d.mx * (Idx)d.my;
    const Idx base = id - (Idx)n * st;   // point with index 0 along DIR
#define NB(tt) N(DIR == 2 ? (tt) : k, DIR == 1 ? (tt) : j, DIR == 0 ? (tt) : i)
#define UB(tt, q) U(DIR == 2 ? (tt) : k, DIR == 1 ? (tt) : j, DIR == 0 ? (tt) : i, q)
    // den: divisor of the 4th-order ucat difference of -skew; next to a body
    // only the k faces reset it (momentum.c:705-723, 967-985, 1202-1219)
    int L = n - 1, R = n + 2;
    double den = 3.;
    if (!PLAIN) {
        if (n == 0 || n == m - 2) {
            if (per) L = m - 3, R = 2;
            else L = n, R = n + 1, den = 1.;
        } else if (NB(L) + NB(R) > 0.1) {
            L = n, R = n + 1;
            if (DIR == 2) den = 1.;
        }
    }
    if (d.second_order) L = n, R = n + 1, den = 1.;
    const S ucon_f = UCon(id);
    S ucon = ucon_f;
    if (!PLAIN) {
        if (per && (n == 0 || n == m - 2)) ucon = UCon((Idx)(base + (Idx)(m - 2) * st));
        // outlet: the last face convects with the flux before it (momentum.c:1227-1229)
        if (DIR == 2 && n == m - 2 && d.bc[5] == 4 && (int)NB(n) == 0) ucon = UCon((Idx)(id - st));
    }
    const S up = -0.5 * (ucon + fabs(ucon));
    const S um = -0.5 * (ucon - fabs(ucon));
    if (!PLAIN && d.immersed && n != m - 2 && NB(n) > 0.1) {
#pragma unroll
        for (int q = 0; q < 3; q++)
            D[q] = um * (0.125 * (-UB(n + 2, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q)) +
                   up * (0.125 * (-UB(n, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q));
    } else if (!PLAIN && d.immersed && n != 0 && NB(n + 1) > 0.1) {
#pragma unroll
        for (int q = 0; q < 3; q++)
            D[q] = um * (0.125 * (-UB(n + 1, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q)) +
                   up * (0.125 * (-UB(n - 1, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q));
    } else if (WENO && d.inviscid) {   // -inv: WENO3 whatever else is set (momentum.c:757-773)
#pragma unroll
        for (int q = 0; q < 3; q++) D[q] = -ucon_f * weno3<S>(UB(L, q), UB(n, q), UB(n + 1, q), UB(R, q), ucon_f);
    } else if (d.second_order) {
#pragma unroll
        for (int q = 0; q < 3; q++) D[q] = -ucon_f * 0.5 * (UB(n, q) + UB(n + 1, q));
    } else if (WENO && d.weno) {   // momentum.c:782-786, 1044-1048, 1283-1287
#pragma unroll
        for (int q = 0; q < 3; q++) D[q] = -ucon_f * weno3<S>(UB(L, q), UB(n, q), UB(n + 1, q), UB(R, q), ucon_f);
    } else {
#pragma unroll
        for (int q = 0; q < 3; q++) D[q] = -ucon_f * 0.0625 * (-UB(L, q) + 9. * UB(n, q) + 9. * UB(n + 1, q) - UB(R, q));
    }
    if (OPTS && d.skew) {   // momentum.c:798-811, 1060-1071, 1298-1309
        const S *dn = DIR == 0 ? dc : DIR == 1 ? de : dz;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            D[q] = D[q] * 0.5;
            const S d4 = (UB(R, q) - UB(L, q)) * (1. / den);
            A[q] = -0.5 * ucon_f * (9. / 8. * dn[q] - 1. / 8. * d4);
        }
    }
    if (!PLAIN && NB(n) + NB(n + 1) > 0.1 && !d.immersed) {
        D[0] = D[1] = D[2] = S(0.0);
        if (OPTS && d.skew) A[0] = A[1] = A[2] = S(0.0);
    }

    const double nu = 1. / d.ren;
    S v0, v1, v2;
    if (d.les) {
        double nut;
        const Idx id1 = id + st;   // the point across the face
        if (!PLAIN && ((n == 0 && !per) || NB(n) > 0.1)) nut = ldx(nu_t, id1);
        else if (!PLAIN && ((n == m - 2 && !per) || NB(n + 1) > 0.1)) nut = ldx(nu_t, id);
        else nut = 0.5 * (ldx(nu_t, id) + ldx(nu_t, id1));
        v0 = (g11 * dc[0] + g21 * de[0] + g31 * dz[0] + r11 * n0 + r21 * n1 + r31 * n2) * ajc * (nut);
        v1 = (g11 * dc[1] + g21 * de[1] + g31 * dz[1] + r12 * n0 + r22 * n1 + r32 * n2) * ajc * (nut);
        v2 = (g11 * dc[2] + g21 * de[2] + g31 * dz[2] + r13 * n0 + r23 * n1 + r33 * n2) * ajc * (nut);
    } else {
        v0 = v1 = v2 = S(0.0);
    }
    v0 += (g11 * dc[0] + g21 * de[0] + g31 * dz[0] + r11 * n0 + r21 * n1 + r31 * n2) * ajc * (nu);
    v1 += (g11 * dc[1] + g21 * de[1] + g31 * dz[1] + r12 * n0 + r22 * n1 + r32 * n2) * ajc * (nu);
    v2 += (g11 * dc[2] + g21 * de[2] + g31 * dz[2] + r13 * n0 + r23 * n1 + r33 * n2) * ajc * (nu);
    if (OPTS && d.clark) {   // momentum.c:907-925, 1142-1160, 1385-1403
        // length scales from the cell metrics of point (k,j,i) and the face's
        // Jacobian, as the reference passes them
        double cc[3], ce[3], cz[3], dx, dy, dzl;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            cc[q] = ldx(g.csi, (Idx)(3 * id + q)); ce[q] = ldx(g.eta, (Idx)(3 * id + q));
            cz[q] = ldx(g.zet, (Idx)(3 * id + q));
        }
        dxdydz(ajc, cc, ce, cz, dx, dy, dzl);
        const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dzl * dzl;
        // Compute_du_dxyz (k-omega.c:597-610): r_ab * ajc
        const S du_dx = r11 * ajc, du_dy = r12 * ajc, du_dz = r13 * ajc;
        const S dv_dx = r21 * ajc, dv_dy = r22 * ajc, dv_dz = r23 * ajc;
        const S dw_dx = r31 * ajc, dw_dy = r32 * ajc, dw_dz = r33 * ajc;
        const S t11 = du_dx * du_dx * dx2 + du_dy * du_dy * dy2 + du_dz * du_dz * dz2;
        const S t12 = du_dx * dv_dx * dx2 + du_dy * dv_dy * dy2 + du_dz * dv_dz * dz2;
        const S t13 = du_dx * dw_dx * dx2 + du_dy * dw_dy * dy2 + du_dz * dw_dz * dz2;
        const S t22 = dv_dx * dv_dx * dx2 + dv_dy * dv_dy * dy2 + dv_dz * dv_dz * dz2;
        const S t23 = dv_dx * dw_dx * dx2 + dv_dy * dw_dy * dy2 + dv_dz * dw_dz * dz2;
        const S t33 = dw_dx * dw_dx * dx2 + dw_dy * dw_dy * dy2 + dw_dz * dw_dz * dz2;
        const S t21 = t12, t31 = t13, t32 = t23;
        v0 = v0 - (t11 * n0 + t12 * n1 + t13 * n2) / 12.;
        v1 = v1 - (t21 * n0 + t22 * n1 + t23 * n2) / 12.;
        v2 = v2 - (t31 * n0 + t32 * n1 + t33 * n2) / 12.;
    }
    V[0] = v0; V[1] = v1; V[2] = v2;
#undef NB
#undef UB
}

struct UConPlain {
    const double *__restrict__ p;
    int dir;
    __device__ __forceinline__ double operator()(long long o) const { return p[3 * o + dir]; }
    __device__ __forceinline__ double operator()(unsigned o) const { return ld32(p, 3 * o + dir); }
};
struct UConDual {
    const double *__restrict__ p, *__restrict__ t;
    int dir;
    __device__ __forceinline__ Dual operator()(long long o) const { return Dual(p[3 * o + dir], t[3 * o + dir]); }
};

// the reference-structured face pass, with -skew and -clark
template <int DIR>
__device__ __forceinline__ void face_flux(const Dims &d, const Geom &g, const double *__restrict__ nvert,
                                          const double *__restrict__ ucont, const double *__restrict__ ucat,
                                          const double *__restrict__ nu_t, int i, int j, int k, double D[3], double V[3],
                                          double A[3])
{
    face_flux_t<DIR, double, GlobalVec, UConPlain, GlobalSca, StoredFaceMetric, true, long long, false, true>(
        d, g, GlobalSca{nvert, d}, UConPlain{ucont, DIR}, GlobalVec{ucat, d}, nu_t, i, j, k, D, V, StoredFaceMetric(), A);
}

// index of a face along a periodic direction as stored: face 0 is a copy of
// face m-2 (momentum.c:1463-1504)
__device__ __forceinline__ int wrap_face(int f, int m, int per) { return (per && f == 0) ? m - 2 : f; }

// Face indices read by the cell divergence of cell (k,j,i): the upwind face
// of the second-order difference (im/jm/km) and the far faces and divisor of
// the 4th-order correction (iR/iL/den_i ...), momentum.c:1506-1600.
struct CellFaces {
    int im, jm, km, iR, iL, jR, jL, kR, kL;
    double den_i, den_j, den_k;
};

template <typename Idx = long long>
__device__ __forceinline__ CellFaces cell_faces(const Dims &d, const double *__restrict__ nvert, int i, int j, int k)
{
    const int ip = d.ip, jp = d.jp, kp = d.kp, mx = d.mx, my = d.my, mz = d.mz;
#define NV(kk, jj, ii) ldx(nvert, ixt<Idx>(d, kk, jj, ii))
    CellFaces c;
    c.im = wrap_face(i - 1, mx, ip); c.jm = wrap_face(j - 1, my, jp); c.km = wrap_face(k - 1, mz, kp);
    int iR = i + 1, iL = i - 2, jR = j + 1, jL = j - 2, kR = k + 1, kL = k - 2;
    double den_i = 3., den_j = 3., den_k = 3.;
    if (i == 1) { if (ip) iL = mx - 3; else iR = i, iL = i - 1, den_i = 1.; }
    else if (i == 2 || i == mx - 3) { if (!ip) iR = i, iL = i - 1, den_i = 1.; }
    else if (i == mx - 2) { if (ip) iR = 1; else iR = i, iL = i - 1, den_i = 1.; }
    if (j == 1) { if (jp) jL = my - 3; else jR = j, jL = j - 1, den_j = 1.; }
    else if (j == 2 || j == my - 3) { if (!jp) jR = j, jL = j - 1, den_j = 1.; }
    else if (j == my - 2) { if (jp) jR = 1; else jR = j, jL = j - 1, den_j = 1.; }
    if (k == 1) { if (kp) kL = mz - 3; else kR = k, kL = k - 1, den_k = 1.; }
    else if (k == 2 || k == mz - 3) { if (!kp) kR = k, kL = k - 1, den_k = 1.; }
    else if (k == mz - 2) { if (kp) kR = 1; else kR = k, kL = k - 1, den_k = 1.; }
    if (NV(k, j, i - 1) + NV(k, j, i) + NV(k, j, i + 1) > 0.1) iR = i, iL = i - 1, den_i = 1.;
    if (NV(k, j - 1, i) + NV(k, j, i) + NV(k, j + 1, i) > 0.1) jR = j, jL = j - 1, den_j = 1.;
    if (NV(k - 1, j, i) + NV(k, j, i) + NV(k + 1, j, i) > 0.1) kR = k, kL = k - 1, den_k = 1.;
    c.iL = wrap_face(iL, mx, ip); c.jL = wrap_face(jL, my, jp); c.kL = wrap_face(kL, mz, kp);
    c.iR = iR; c.jR = jR; c.kR = kR;
    c.den_i = den_i; c.den_j = den_j; c.den_k = den_k;
#undef NV
    return c;
}

// fp of cell (k,j,i) from the face arrays (momentum.c:1506-1636).
// face: slot-major SoA, slot s = 3*kind + comp with kind 0..2 div1..3, 3..5
// visc1..3 and, under -skew, 6..8 adv1..3.
__device__ __forceinline__ void cell_fp(const Dims &d, const double *__restrict__ nvert,
                                        const double *__restrict__ face, int i, int j, int k, double f[3])
{
    const long long P = d.P;
#define FI(kind, q, kk, jj, ii) face[(long long)(3 * (kind) + (q)) * P + IX(d, kk, jj, ii)]
    const CellFaces c = cell_faces(d, nvert, i, j, k);
    const int im = c.im, jm = c.jm, km = c.km;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        double dv = (FI(0, q, k, j, i) - FI(0, q, k, j, im) + FI(1, q, k, j, i) - FI(1, q, k, jm, i) +
                     FI(2, q, k, j, i) - FI(2, q, km, j, i));
        double r;
        if (d.second_order || d.inviscid) {   // momentum.c:1582-1588
            r = dv;
        } else {
            double d4 = 0;
            d4 += (FI(0, q, k, j, c.iR) - FI(0, q, k, j, c.iL)) * (1. / c.den_i);
            d4 += (FI(1, q, k, c.jR, i) - FI(1, q, k, c.jL, i)) * (1. / c.den_j);
            d4 += (FI(2, q, c.kR, j, i) - FI(2, q, c.kL, j, i)) * (1. / c.den_k);
            r = 9. / 8. * dv + -1. / 8. * d4;
        }
        if (d.skew && !d.inviscid) {   // momentum.c:1596-1608: the advective fluxes, 4th-order interpolated
            int iR = c.iR, iL = c.iL, jR = c.jR, jL = c.jL, kR = c.kR, kL = c.kL;
            if (d.second_order) iR = i, iL = im, jR = j, jL = jm, kR = k, kL = km;
            r += 9. / 8. * 0.5 * (FI(6, q, k, j, i) + FI(6, q, k, j, im)) - 1. / 8. * 0.5 * (FI(6, q, k, j, iR) + FI(6, q, k, j, iL));
            r += 9. / 8. * 0.5 * (FI(7, q, k, j, i) + FI(7, q, k, jm, i)) - 1. / 8. * 0.5 * (FI(7, q, k, jR, i) + FI(7, q, k, jL, i));
            r += 9. / 8. * 0.5 * (FI(8, q, k, j, i) + FI(8, q, km, j, i)) - 1. / 8. * 0.5 * (FI(8, q, kR, j, i) + FI(8, q, kL, j, i));
        }
        if (!d.inviscid)   // momentum.c:1621-1627
            r += (FI(3, q, k, j, i) - FI(3, q, k, j, im) + FI(4, q, k, j, i) - FI(4, q, k, jm, i) +
                  FI(5, q, k, j, i) - FI(5, q, km, j, i)) / 1.0;
        f[q] = r;
    }
#undef FI
}

// cell index of a periodic ghost read: cell m-1 is a copy of cell 1
__device__ __forceinline__ int wrap_cell(int c, int m, int per) { return (per && c == m - 1) ? 1 : c; }

// rhs += scale * (contravariant projection of fp) with the face masks
// (momentum.c:1673-1693, 1795-1803), from fp values already at hand: fc at
// the cell, fa/fb/fcc at its i+1 / j+1 / k+1 neighbours (periodic-wrapped).
__device__ __forceinline__ void project_rhs_v(const Dims &d, const Geom &g, const double *__restrict__ nvert, int i,
                                              int j, int k, long long id, double scale, const double fc[3],
                                              const double fa[3], const double fb[3], const double fcc[3], double r[3])
{
    const long long ga = id + 1, gb = id + d.mx, gc = id + (long long)d.mx * d.my;  // metric ghosts are stored
    const double *cs = g.csi, *et = g.eta, *ze = g.zet;
    r[0] += scale * (0.5 * (cs[3 * id] * fc[0] + cs[3 * id + 1] * fc[1] + cs[3 * id + 2] * fc[2]) +
                     0.5 * (cs[3 * ga] * fa[0] + cs[3 * ga + 1] * fa[1] + cs[3 * ga + 2] * fa[2])) * g.iaj[id];
    r[1] += scale * (0.5 * (et[3 * id] * fc[0] + et[3 * id + 1] * fc[1] + et[3 * id + 2] * fc[2]) +
                     0.5 * (et[3 * gb] * fb[0] + et[3 * gb + 1] * fb[1] + et[3 * gb + 2] * fb[2])) * g.jaj[id];
    r[2] += scale * (0.5 * (ze[3 * id] * fc[0] + ze[3 * id + 1] * fc[1] + ze[3 * id + 2] * fc[2]) +
                     0.5 * (ze[3 * gc] * fcc[0] + ze[3 * gc + 1] * fcc[1] + ze[3 * gc + 2] * fcc[2])) * g.kaj[id];
    const double n0 = nvert[id];
    if (n0 + nvert[id + 1] > 0.1 || (!d.ip && i == d.mx - 2)) r[0] = 0;
    if (n0 + nvert[id + d.mx] > 0.1 || (!d.jp && j == d.my - 2)) r[1] = 0;
    if (n0 + nvert[id + (long long)d.mx * d.my] > 0.1 || (!d.kp && k == d.mz - 2)) r[2] = 0;
}

// rhs += scale * (contravariant projection of fp) with the face masks
// (momentum.c:1673-1693, 1795-1803)
template <typename Idx = long long>
__device__ __forceinline__ void project_rhs(const Dims &d, const Geom &g, const double *__restrict__ nvert,
                                            const double *__restrict__ fp, int i, int j, int k, Idx id,
                                            double scale, double r[3])
{
    const Idx a = ixt<Idx>(d, k, j, wrap_cell(i + 1, d.mx, d.ip)), b = ixt<Idx>(d, k, wrap_cell(j + 1, d.my, d.jp), i),
              c = ixt<Idx>(d, wrap_cell(k + 1, d.mz, d.kp), j, i);
    const Idx sy = (Idx)d.mx, sz = (Idx)d.mx * (Idx)d.my;
    const Idx ga = id + 1, gb = id + sy, gc = id + sz;  // metric ghosts are stored
    const double *cs = g.csi, *et = g.eta, *ze = g.zet;
#define L3(p, o, q) ldx(p, (Idx)(3 * (o) + (q)))
    r[0] += scale * (0.5 * (L3(cs, id, 0) * L3(fp, id, 0) + L3(cs, id, 1) * L3(fp, id, 1) + L3(cs, id, 2) * L3(fp, id, 2)) +
                     0.5 * (L3(cs, ga, 0) * L3(fp, a, 0) + L3(cs, ga, 1) * L3(fp, a, 1) + L3(cs, ga, 2) * L3(fp, a, 2))) *
            ldx(g.iaj, id);
    r[1] += scale * (0.5 * (L3(et, id, 0) * L3(fp, id, 0) + L3(et, id, 1) * L3(fp, id, 1) + L3(et, id, 2) * L3(fp, id, 2)) +
                     0.5 * (L3(et, gb, 0) * L3(fp, b, 0) + L3(et, gb, 1) * L3(fp, b, 1) + L3(et, gb, 2) * L3(fp, b, 2))) *
            ldx(g.jaj, id);
    r[2] += scale * (0.5 * (L3(ze, id, 0) * L3(fp, id, 0) + L3(ze, id, 1) * L3(fp, id, 1) + L3(ze, id, 2) * L3(fp, id, 2)) +
                     0.5 * (L3(ze, gc, 0) * L3(fp, c, 0) + L3(ze, gc, 1) * L3(fp, c, 1) + L3(ze, gc, 2) * L3(fp, c, 2))) *
            ldx(g.kaj, id);
#undef L3
    const double n0 = ldx(nvert, id);
    if (n0 + ldx(nvert, (Idx)(id + 1)) > 0.1 || (!d.ip && i == d.mx - 2)) r[0] = 0;
    if (n0 + ldx(nvert, (Idx)(id + sy)) > 0.1 || (!d.jp && j == d.my - 2)) r[1] = 0;
    if (n0 + ldx(nvert, (Idx)(id + sz)) > 0.1 || (!d.kp && k == d.mz - 2)) r[2] = 0;
}
