// Transposes of the per-point stencil bodies of the momentum right-hand side.
// Each function mirrors the branch structure of its forward counterpart in
// vfs_stencil.h (cited there against momentum.c / k-omega.c) and scatters
// the adjoint of its outputs into the adjoints of the inputs it read.
//
// The adjoint uses hardware FP64 atomics on device memory from hipMalloc
// (coarse grained): sums are exact up to the order of additions.
#pragma once
#include "vfs_stencil.h"

__device__ __forceinline__ void adj_add(double *p, double v)
{
    if (v != 0.0) unsafeAtomicAdd(p, v);
}

// Scatter into the adjoint of a cell-centred vector field at (k,j,i).
struct VecAdj {
    double *p;
    Dims d;
    __device__ __forceinline__ void operator()(int k, int j, int i, int q, double v) const
    {
        adj_add(p + 3 * IX(d, k, j, i) + q, v);
    }
};

// Transpose of du_face: dcb/deb/dzb are the adjoints of dc/de/dz.
template <typename Acc>
__device__ __forceinline__ void du_face_adj(int dir, const Acc &A, const double *__restrict__ nvert, const Dims &d,
                                            int i, int j, int k, const double dcb[3], const double deb[3],
                                            const double dzb[3])
{
#define NV(kk, jj, ii) nvert[IX(d, kk, jj, ii)]
// (U(a) + U(b) - U(c) - U(e)) * s
#define T4(s, bar, ka, ja, ia, kb, jb, ib, kc, jc, ic, ke, je, ie)                 \
    for (int q = 0; q < 3; q++) {                                                  \
        double w = (bar)[q] * (s);                                                 \
        A(ka, ja, ia, q, w); A(kb, jb, ib, q, w);                                  \
        A(kc, jc, ic, q, -w); A(ke, je, ie, q, -w);                                \
    }
// U(a) - U(b)
#define T2(bar, ka, ja, ia, kb, jb, ib)                                            \
    for (int q = 0; q < 3; q++) { A(ka, ja, ia, q, (bar)[q]); A(kb, jb, ib, q, -(bar)[q]); }
    if (dir == 0) {
        T2(dcb, k, j, i + 1, k, j, i);
        if (NV(k, j + 1, i) > VFS_SOLID || NV(k, j + 1, i + 1) > VFS_SOLID) {
            T4(0.5, deb, k, j, i + 1, k, j, i, k, j - 1, i + 1, k, j - 1, i);
        } else if (NV(k, j - 1, i) > VFS_SOLID || NV(k, j - 1, i + 1) > VFS_SOLID) {
            T4(0.5, deb, k, j + 1, i + 1, k, j + 1, i, k, j, i + 1, k, j, i);
        } else {
            T4(0.25, deb, k, j + 1, i + 1, k, j + 1, i, k, j - 1, i + 1, k, j - 1, i);
        }
        if (NV(k + 1, j, i) > VFS_SOLID || NV(k + 1, j, i + 1) > VFS_SOLID) {
            T4(0.5, dzb, k, j, i + 1, k, j, i, k - 1, j, i + 1, k - 1, j, i);
        } else if (NV(k - 1, j, i) > VFS_SOLID || NV(k - 1, j, i + 1) > VFS_SOLID) {
            T4(0.5, dzb, k + 1, j, i + 1, k + 1, j, i, k, j, i + 1, k, j, i);
        } else {
            T4(0.25, dzb, k + 1, j, i + 1, k + 1, j, i, k - 1, j, i + 1, k - 1, j, i);
        }
    } else if (dir == 1) {
        if (NV(k, j, i + 1) > VFS_SOLID || NV(k, j + 1, i + 1) > VFS_SOLID) {
            T4(0.5, dcb, k, j + 1, i, k, j, i, k, j + 1, i - 1, k, j, i - 1);
        } else if (NV(k, j, i - 1) > VFS_SOLID || NV(k, j + 1, i - 1) > VFS_SOLID) {
            T4(0.5, dcb, k, j + 1, i + 1, k, j, i + 1, k, j + 1, i, k, j, i);
        } else {
            T4(0.25, dcb, k, j + 1, i + 1, k, j, i + 1, k, j + 1, i - 1, k, j, i - 1);
        }
        T2(deb, k, j + 1, i, k, j, i);
        if (NV(k + 1, j, i) > VFS_SOLID || NV(k + 1, j + 1, i) > VFS_SOLID) {
            T4(0.5, dzb, k, j + 1, i, k, j, i, k - 1, j + 1, i, k - 1, j, i);
        } else if (NV(k - 1, j, i) > VFS_SOLID || NV(k - 1, j + 1, i) > VFS_SOLID) {
            T4(0.5, dzb, k + 1, j + 1, i, k + 1, j, i, k, j + 1, i, k, j, i);
        } else {
            T4(0.25, dzb, k + 1, j + 1, i, k + 1, j, i, k - 1, j + 1, i, k - 1, j, i);
        }
    } else {
        if (NV(k, j, i + 1) > VFS_SOLID || NV(k + 1, j, i + 1) > VFS_SOLID) {
            T4(0.5, dcb, k + 1, j, i, k, j, i, k + 1, j, i - 1, k, j, i - 1);
        } else if (NV(k, j, i - 1) > VFS_SOLID || NV(k + 1, j, i - 1) > VFS_SOLID) {
            T4(0.5, dcb, k + 1, j, i + 1, k, j, i + 1, k + 1, j, i, k, j, i);
        } else {
            T4(0.25, dcb, k + 1, j, i + 1, k, j, i + 1, k + 1, j, i - 1, k, j, i - 1);
        }
        if (NV(k, j + 1, i) > VFS_SOLID || NV(k + 1, j + 1, i) > VFS_SOLID) {
            T4(0.5, deb, k + 1, j, i, k, j, i, k + 1, j - 1, i, k, j - 1, i);
        } else if (NV(k, j - 1, i) > VFS_SOLID || NV(k + 1, j - 1, i) > VFS_SOLID) {
            T4(0.5, deb, k + 1, j + 1, i, k, j + 1, i, k + 1, j, i, k, j, i);
        } else {
            T4(0.25, deb, k + 1, j + 1, i, k, j + 1, i, k + 1, j - 1, i, k, j - 1, i);
        }
        T2(dzb, k + 1, j, i, k, j, i);
    }
#undef T4
#undef T2
#undef NV
}

// weno3 (level.c:1734-1771) and its gradient with respect to f0..f3 (the
// wave speed only selects the upwind side).  With fL, fC, fR the upwind
// triple, w0 = alpha0 / (alpha0 + alpha1), W = w0 u0 + (1 - w0) u1:
//   dW = (u0 - u1) dw0 + w0 du0 + w1 du1,
//   dw0 = (alpha1 dalpha0 - alpha0 dalpha1) / (alpha0 + alpha1)^2,
//   dalpha_m = -2 d_m / (eps + beta_m)^3 dbeta_m.
__device__ __forceinline__ double weno3_grad(double f0, double f1, double f2, double f3, double ws, double g[4])
{
    const bool mir = !(ws > 0);
    const double fL = mir ? f3 : f0, fC = mir ? f2 : f1, fR = mir ? f1 : f2;
    const double d0 = 2. / 3., d1 = 1. / 3., eps = 1.e-6;
    const double b0 = fC - fR, b1 = fL - fC;
    const double e0 = eps + b0 * b0, e1 = eps + b1 * b1;
    const double a0 = d0 / (e0 * e0), a1 = d1 / (e1 * e1);
    const double sa = a0 + a1, w0 = a0 / sa, w1 = a1 / sa;
    const double u0 = fC * 0.5 + fR * 0.5, u1 = -fL * 0.5 + fC * 1.5;
    const double da0 = -2. * d0 / (e0 * e0 * e0), da1 = -2. * d1 / (e1 * e1 * e1);   // d alpha / d beta
    const double A = (u0 - u1) / (sa * sa);
    // d beta0 = 2 b0 (dfC - dfR), d beta1 = 2 b1 (dfL - dfC)
    const double gL = A * (-a0 * da1 * 2. * b1) - 0.5 * w1;
    const double gC = A * (a1 * da0 * 2. * b0 + a0 * da1 * 2. * b1) + 0.5 * w0 + 1.5 * w1;
    const double gR = A * (-a1 * da0 * 2. * b0) + 0.5 * w0;
    if (!mir) { g[0] = gL; g[1] = gC; g[2] = gR; g[3] = 0; }
    else { g[3] = gL; g[2] = gC; g[1] = gR; g[0] = 0; }
    return w0 * u0 + w1 * u1;
}

// Transpose of face_flux_t<DIR, double> with respect to ucont (component DIR)
// and ucat, at the linearisation point (ucont, ucat); nu_t is frozen.
// Db/Vb are the adjoints of the convective and viscous face fluxes.
template <int DIR>
__device__ __forceinline__ void face_flux_adj(const Dims &d, const Geom &g, const double *__restrict__ nvert,
                                              const double *__restrict__ ucont, const double *__restrict__ ucat,
                                              const double *__restrict__ nu_t, int i, int j, int k,
                                              const double Db[3], const double Vb[3], double *ucont_b, double *ucat_b)
{
    const long long id = IX(d, k, j, i);
    const double *fc = DIR == 0 ? g.icsi : DIR == 1 ? g.jcsi : g.kcsi;
    const double *fe = DIR == 0 ? g.ieta : DIR == 1 ? g.jeta : g.keta;
    const double *fz = DIR == 0 ? g.izet : DIR == 1 ? g.jzet : g.kzet;
    const double *faj = DIR == 0 ? g.iaj : DIR == 1 ? g.jaj : g.kaj;
    const double ajc = faj[id];
    const double cs[3] = {fc[3 * id], fc[3 * id + 1], fc[3 * id + 2]};
    const double et[3] = {fe[3 * id], fe[3 * id + 1], fe[3 * id + 2]};
    const double ze[3] = {fz[3 * id], fz[3 * id + 1], fz[3 * id + 2]};
    const double *nn = DIR == 0 ? cs : DIR == 1 ? et : ze;
    const double g11 = cs[0] * nn[0] + cs[1] * nn[1] + cs[2] * nn[2];
    const double g21 = et[0] * nn[0] + et[1] * nn[1] + et[2] * nn[2];
    const double g31 = ze[0] * nn[0] + ze[1] * nn[1] + ze[2] * nn[2];

    const int n = DIR == 0 ? i : DIR == 1 ? j : k;
    const int m = DIR == 0 ? d.mx : DIR == 1 ? d.my : d.mz;
    const int per = DIR == 0 ? d.ip : DIR == 1 ? d.jp : d.kp;
    const long long st = DIR == 0 ? 1 : DIR == 1 ? (long long)d.mx : (long long)d.mx * d.my;
    const long long base = id - (long long)n * st;
#define NB(tt) nvert[base + (long long)(tt) * st]
#define UB(tt, q) ucat[3 * (base + (long long)(tt) * st) + (q)]
#define AU(tt, q, v) adj_add(ucat_b + 3 * (base + (long long)(tt) * st) + (q), (v))
    int L = n - 1, R = n + 2;
    if (n == 0 || n == m - 2) {
        if (per) L = m - 3, R = 2;
        else L = n, R = n + 1;
    } else if (NB(L) + NB(R) > 0.1) {
        L = n, R = n + 1;
    }
    if (d.second_order) L = n, R = n + 1;
    const long long of = base + (long long)n * st;
    const long long ou = (per && (n == 0 || n == m - 2)) ? base + (long long)(m - 2) * st : of;
    const double ucon_f = ucont[3 * of + DIR];
    const double ucon = ucont[3 * ou + DIR];
    const double sg = ucon > 0 ? 1.0 : (ucon < 0 ? -1.0 : 0.0);
    const double up = -0.5 * (ucon + fabs(ucon));
    const double um = -0.5 * (ucon - fabs(ucon));
    double ucon_fb = 0, uconb = 0;
    const bool masked = NB(n) + NB(n + 1) > 0.1 && !d.immersed;
    if (!masked) {
        if (d.immersed && n != m - 2 && NB(n) > 0.1) {
            double umb = 0, upb = 0;
            for (int q = 0; q < 3; q++) {
                const double a = 0.125 * (-UB(n + 2, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q);
                const double b = 0.125 * (-UB(n, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q);
                umb += Db[q] * a;
                upb += Db[q] * b;
                AU(n + 2, q, Db[q] * um * -0.125);
                AU(n + 1, q, Db[q] * (um * 0.75 + up * 0.375));
                AU(n, q, Db[q] * (um * 0.375 + up * 0.625));
            }
            uconb += upb * (-0.5 * (1. + sg)) + umb * (-0.5 * (1. - sg));
        } else if (d.immersed && n != 0 && NB(n + 1) > 0.1) {
            double umb = 0, upb = 0;
            for (int q = 0; q < 3; q++) {
                const double a = 0.125 * (-UB(n + 1, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q);
                const double b = 0.125 * (-UB(n - 1, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q);
                umb += Db[q] * a;
                upb += Db[q] * b;
                AU(n + 1, q, Db[q] * (um * 0.625 + up * 0.375));
                AU(n, q, Db[q] * (um * 0.375 + up * 0.75));
                AU(n - 1, q, Db[q] * up * -0.125);
            }
            uconb += upb * (-0.5 * (1. + sg)) + umb * (-0.5 * (1. - sg));
        } else if (d.second_order) {
            for (int q = 0; q < 3; q++) {
                ucon_fb += Db[q] * (-0.5 * (UB(n, q) + UB(n + 1, q)));
                AU(n, q, Db[q] * -ucon_f * 0.5);
                AU(n + 1, q, Db[q] * -ucon_f * 0.5);
            }
        } else if (d.weno) {
            for (int q = 0; q < 3; q++) {
                double gw[4];
                const double W = weno3_grad(UB(L, q), UB(n, q), UB(n + 1, q), UB(R, q), ucon_f, gw);
                ucon_fb += Db[q] * -W;
                const double c = Db[q] * -ucon_f;
                AU(L, q, c * gw[0]);
                AU(n, q, c * gw[1]);
                AU(n + 1, q, c * gw[2]);
                AU(R, q, c * gw[3]);
            }
        } else {
            const double c = -ucon_f * 0.0625;
            for (int q = 0; q < 3; q++) {
                ucon_fb += Db[q] * (-0.0625 * (-UB(L, q) + 9. * UB(n, q) + 9. * UB(n + 1, q) - UB(R, q)));
                AU(L, q, -c * Db[q]);
                AU(n, q, 9. * c * Db[q]);
                AU(n + 1, q, 9. * c * Db[q]);
                AU(R, q, -c * Db[q]);
            }
        }
        adj_add(ucont_b + 3 * of + DIR, ucon_fb);
        adj_add(ucont_b + 3 * ou + DIR, uconb);
    }

    double W = ajc * (1. / d.ren);
    if (d.les) {
        double nut;
        if ((n == 0 && !per) || NB(n) > 0.1) nut = nu_t[base + (long long)(n + 1) * st];
        else if ((n == m - 2 && !per) || NB(n + 1) > 0.1) nut = nu_t[id];
        else nut = 0.5 * (nu_t[id] + nu_t[base + (long long)(n + 1) * st]);
        W += ajc * nut;
    }
    double dcb[3] = {0, 0, 0}, deb[3] = {0, 0, 0}, dzb[3] = {0, 0, 0};
    for (int q = 0; q < 3; q++) {
        const double vb = Vb[q] * W;
        dcb[q] += vb * g11;
        deb[q] += vb * g21;
        dzb[q] += vb * g31;
        for (int a = 0; a < 3; a++) {
            dcb[a] += vb * nn[a] * cs[q];
            deb[a] += vb * nn[a] * et[q];
            dzb[a] += vb * nn[a] * ze[q];
        }
    }
    du_face_adj(DIR, VecAdj{ucat_b, d}, nvert, d, i, j, k, dcb, deb, dzb);
#undef NB
#undef UB
#undef AU
}
