// Transposes of the per-point stencil bodies of the momentum right-hand side.
// Each function mirrors the branch structure of its forward counterpart in
// vfs_stencil.h (cited there against momentum.c / k-omega.c).
//
// The transpose of a face flux does not scatter into the adjoints of the
// points it read.  It returns the adjoint per stencil position of the face
// (FaceAdj), and every point later gathers the positions that name it from
// the faces around it in a fixed order, so J^T v is the same bits on every
// run.
#pragma once
#include "vfs_stencil.h"

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

// The adjoint of one face flux, by stencil position.  For a face of stored
// index n along DIR (cells c0 = n and c1 = n+1 along DIR):
//  S[r][q]     ucat component q at the face's own row, along-DIR position
//              n + r - 1 (r = 0..3); in a periodic direction the first face
//              (n = 0) has r = 0 at m-3 and the last one (n = m-2) has r = 3
//              at 2, the positions its 4th-order stencil wraps to;
//  X[tt][s][q] ucat component q one row to the - (s = 0) or + (s = 1) side in
//              the tt-th transverse direction, at positions n and n+1 alike;
//  uself       ucont component DIR of the face itself;
//  usrc        ucont component DIR of face m-2, which the first face of a
//              periodic direction convects with.
struct FaceAdj {
    double S[4][3];
    double X[2][2][3];
    double uself, usrc;
};
constexpr int FACE_ADJ_SLOTS = 26;

// Transpose of face_flux_t<DIR, double> with respect to ucont (component DIR)
// and ucat, at the linearisation point (ucont, ucat); nu_t is frozen.
// Db/Vb are the adjoints of the convective and viscous face fluxes.
template <int DIR>
__device__ __forceinline__ void face_flux_adj(const Dims &d, const Geom &g, const double *__restrict__ nvert,
                                              const double *__restrict__ ucont, const double *__restrict__ ucat,
                                              const double *__restrict__ nu_t, int i, int j, int k,
                                              const double Db[3], const double Vb[3], FaceAdj &F)
{
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 3; q++) F.S[r][q] = 0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int q = 0; q < 3; q++) F.X[a][s][q] = 0;
    F.uself = F.usrc = 0;
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
    int L = n - 1, R = n + 2;
    if (n == 0 || n == m - 2) {
        if (per) L = m - 3, R = 2;
        else L = n, R = n + 1;
    } else if (NB(L) + NB(R) > 0.1) {
        L = n, R = n + 1;
    }
    if (d.second_order) L = n, R = n + 1;
    // stencil slots of L and R: the upwind neighbours, or the face's own cells
    const int sL = L == n ? 1 : 0, sR = R == n + 1 ? 2 : 3;
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
                F.S[3][q] += Db[q] * um * -0.125;
                F.S[2][q] += Db[q] * (um * 0.75 + up * 0.375);
                F.S[1][q] += Db[q] * (um * 0.375 + up * 0.625);
            }
            uconb += upb * (-0.5 * (1. + sg)) + umb * (-0.5 * (1. - sg));
        } else if (d.immersed && n != 0 && NB(n + 1) > 0.1) {
            double umb = 0, upb = 0;
            for (int q = 0; q < 3; q++) {
                const double a = 0.125 * (-UB(n + 1, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q);
                const double b = 0.125 * (-UB(n - 1, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q);
                umb += Db[q] * a;
                upb += Db[q] * b;
                F.S[2][q] += Db[q] * (um * 0.625 + up * 0.375);
                F.S[1][q] += Db[q] * (um * 0.375 + up * 0.75);
                F.S[0][q] += Db[q] * up * -0.125;
            }
            uconb += upb * (-0.5 * (1. + sg)) + umb * (-0.5 * (1. - sg));
        } else if (d.second_order) {
            for (int q = 0; q < 3; q++) {
                ucon_fb += Db[q] * (-0.5 * (UB(n, q) + UB(n + 1, q)));
                F.S[1][q] += Db[q] * -ucon_f * 0.5;
                F.S[2][q] += Db[q] * -ucon_f * 0.5;
            }
        } else if (d.weno) {
            for (int q = 0; q < 3; q++) {
                double gw[4];
                const double W = weno3_grad(UB(L, q), UB(n, q), UB(n + 1, q), UB(R, q), ucon_f, gw);
                ucon_fb += Db[q] * -W;
                const double c = Db[q] * -ucon_f;
                F.S[sL][q] += c * gw[0];
                F.S[1][q] += c * gw[1];
                F.S[2][q] += c * gw[2];
                F.S[sR][q] += c * gw[3];
            }
        } else {
            const double c = -ucon_f * 0.0625;
            for (int q = 0; q < 3; q++) {
                ucon_fb += Db[q] * (-0.0625 * (-UB(L, q) + 9. * UB(n, q) + 9. * UB(n + 1, q) - UB(R, q)));
                F.S[sL][q] += -c * Db[q];
                F.S[1][q] += 9. * c * Db[q];
                F.S[2][q] += 9. * c * Db[q];
                F.S[sR][q] += -c * Db[q];
            }
        }
        if (ou == of) {
            F.uself = ucon_fb + uconb;
        } else {
            F.uself = ucon_fb;
            F.usrc = uconb;
        }
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
    // transpose of du_face: the normal difference U(c1) - U(c0), then the two
    // transverse differences (U(c1 + p) + U(c0 + p) - U(c1 + m) - U(c0 + m)) * s
    // with du_face's choice of (p, m, s)
    const double *dnb = DIR == 0 ? dcb : DIR == 1 ? deb : dzb;
#pragma unroll
    for (int q = 0; q < 3; q++) { F.S[2][q] += dnb[q]; F.S[1][q] -= dnb[q]; }
    const int i1 = i + (DIR == 0), j1 = j + (DIR == 1), k1 = k + (DIR == 2);
#define NV(kk, jj, ii) nvert[IX(d, kk, jj, ii)]
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
        const int t = DIR == 0 ? 1 + tt : DIR == 1 ? 2 * tt : tt;
        const int ti = t == 0, tj = t == 1, tk = t == 2;
        int p = 1, mm = -1;
        double sc = 0.25;
        if (NV(k + tk, j + tj, i + ti) > VFS_SOLID || NV(k1 + tk, j1 + tj, i1 + ti) > VFS_SOLID) {
            p = 0; mm = -1; sc = 0.5;
        } else if (NV(k - tk, j - tj, i - ti) > VFS_SOLID || NV(k1 - tk, j1 - tj, i1 - ti) > VFS_SOLID) {
            p = 1; mm = 0; sc = 0.5;
        }
        const double *tb = t == 0 ? dcb : t == 1 ? deb : dzb;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double w = tb[q] * sc;
            // + w at transverse offset p, - w at offset mm, at c1 and c0 alike
            if (p == 0) { F.S[2][q] += w; F.S[1][q] += w; }
            else F.X[tt][1][q] += w;
            if (mm == 0) { F.S[2][q] -= w; F.S[1][q] -= w; }
            else F.X[tt][0][q] -= w;
        }
    }
#undef NV
#undef NB
#undef UB
}
