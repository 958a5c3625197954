// The dynamic Smagorinsky procedure: Compute_Smagorinsky_Constant_1
// (les.c:67-1133) with its options: -les 1 / 2 / 3 and the fresh-start rule,
// the homogeneous-direction averages, -testfilter_ik and, on a -clark
// context, the Clark model's Germano term.
//
// Five kernels over the padded grid, each followed by the periodic ghost
// copies the reference makes at that point:
//   0. k_dsm_point  point-wise inputs of the filters as unit-stride planes;
//   1. k_dsm_grad   velocity gradient at the non-solid interior cells,
//                   stored as the six products S_ab |S| the test filter of
//                   the second pass reads (the reference keeps the nine
//                   gradients and |S| and forms these products there), the
//                   test-filtered velocity and, under -clark, the nine
//                   gradients;
//   2. k_dsm_lm     the Germano identity in the reference's contravariant
//                   form: LM = L_ba M_aq G_bq (- N_ab M_aq G_bq under
//                   -clark), MM = M_nm M_nl G_ml;
//   3. k_dsm_cs     Cs = max(0.5 <LM> / (<MM> + 1e-4), 0) and the clipping;
//   4. k_dsm_smooth the box filter of the clipped Cs at the fluid cells.
// Scratch: 19 doubles a point (28 under -clark), allocated on first use and
// kept by the context.  Built without FMA contraction; every sum keeps the reference's
// order, so results differ from the oracle only through pow(x, 1/3).
#include "vfs_internal.h"
#include "vfs_stencil.h"

#define VFS_LES_EPS 1.e-4   // les.c:19
#define VFS_WALL_CS 0.001   // les.c:20
#define PMAX(a, b) ((a) < (b) ? (b) : (a))   // PetscMax
#define PMIN(a, b) ((a) < (b) ? (a) : (b))   // PetscMin

// rhs2.c:412-475 integrate_testfilter: the weighted mean over the eight
// 2x2x2 octants of the 3x3x3 box, sum(v1..v8) / sum(w1..w8), each octant
// sum v_o = sum of its eight v * w products in the reference's order.
// Within an octant that order is lexicographic in (j, i, k) offsets, so the
// 27 points are streamed once in that order (OCT_FOR) and each product is
// added to the accumulators of the octants holding it: the same sums,
// rounding for rounding, with 8 accumulators a quantity instead of 27
// stored values.  Octant o holds the k offsets (o & 1) + {0,1}, the i offsets
// ((o >> 1) & 1) + {0,1} and the j offsets ((o >> 2) & 1) + {0,1}.
struct Oct {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
// add product x of the point with k, j, i offsets + 1 = (R, Q, P); with
// run-time offsets the octant tests become selects
__device__ __forceinline__ void oct_add(Oct &a, int R, int Q, int P, double x)
{
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int oa = o & 1, oc = (o >> 1) & 1, ob = (o >> 2) & 1;
        const bool in = (unsigned)(R - oa) <= 1u && (unsigned)(Q - ob) <= 1u && (unsigned)(P - oc) <= 1u;
        const bool first = R == oa && Q == ob && P == oc;
        a.s[o] = in ? (first ? x : a.s[o] + x) : a.s[o];
    }
}
__device__ __forceinline__ double oct_total(const Oct &a)
{
    double t = a.s[0];
#pragma unroll
    for (int o = 1; o < 8; o++) t += a.s[o];
    return t;
}
// rhs2.c:388-410 integrate_testfilter_ik: Simpson's rule over the i-k plane
// through the centre (weights 1, 4, 16 over 36, no volumes); f(R, P) is the
// value at k, i offsets + 1 on the centre's j
template <typename F>
__device__ __forceinline__ double simpson_ik(F f)
{
    return ((f(0, 0) + f(2, 0) + f(0, 2) + f(2, 2)) + 4. * (f(0, 1) + f(1, 0) + f(2, 1) + f(1, 2)) + 16. * f(1, 1)) /
           36.;
}

// sum(v1..v8) of 27 stored products p[9 R + 3 Q + P], in the reference's
// order (the octant's eight terms with the k offset fastest, then i, then j)
__device__ __forceinline__ double octsum27(const double p[27])
{
    double vs = 0;
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int oa = o & 1, oc = (o >> 1) & 1, ob = (o >> 2) & 1;
        double v = 0;
#pragma unroll
        for (int tt = 0; tt < 8; tt++) {
            const int r = 9 * (oa + (tt & 1)) + 3 * (ob + ((tt >> 2) & 1)) + oc + ((tt >> 1) & 1);
            v = tt ? v + p[r] : p[r];
        }
        vs = o ? vs + v : v;
    }
    return vs;
}

// the 27 points in (j, i, k) lexicographic order; R, Q, P are the k, j, i
// offsets + 1.  OCT_FOR unrolls the box; OCT_FOR_RT keeps the nine-point j
// rows as a run-time loop, which bounds the loads the compiler issues ahead.
#define OCT_FOR                                                                                       \
    _Pragma("unroll") for (int Q = 0; Q < 3; Q++) _Pragma("unroll") for (int P = 0; P < 3; P++)       \
        _Pragma("unroll") for (int R = 0; R < 3; R++)
#define OCT_FOR_RT                                                                                    \
    _Pragma("unroll 1") for (int Q = 0; Q < 3; Q++) _Pragma("unroll") for (int P = 0; P < 3; P++)     \
        _Pragma("unroll") for (int R = 0; R < 3; R++)
__device__ __forceinline__ int slot_off(const Dims &d, int R, int Q, int P)
{
    return ((R - 1) * d.my + (Q - 1)) * d.mx + (P - 1);
}

// Point-wise inputs of the box filters, as planes of P doubles so that the
// stencil passes read them with unit stride: the get_weight value (les.c:
// 23-64: 1/aj, 0 in solid cells; a function of the neighbour alone), the
// velocity components and U_b = u . (csi, eta, zet)_b (les.c:348-354).
__global__ void k_dsm_point(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucat,
                            double *__restrict__ wt, double *__restrict__ uq, double *__restrict__ Ub)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)d.P) return;
    const size_t P = (size_t)d.P;
    st32(wt, t, ld32(nvert, t) > 1.1 ? 0 : 1. / ld32(g.aj, t));
    const double u0 = ld32(ucat, 3 * t), u1 = ld32(ucat, 3 * t + 1), u2 = ld32(ucat, 3 * t + 2);
    st32(uq, t, u0);
    st32(uq + P, t, u1);
    st32(uq + 2 * P, t, u2);
    const double *met[3] = {g.csi, g.eta, g.zet};
#pragma unroll
    for (int b = 0; b < 3; b++)
        st32(Ub + b * P, t, u0 * ld32(met[b], 3 * t) + u1 * ld32(met[b], 3 * t + 1) + u2 * ld32(met[b], 3 * t + 2));
}

// a cell whose box holds no ghost point: the Cs-averaging weights are the
// get_weight values and no index wraps
__device__ __forceinline__ bool away_from_ghosts(const Dims &d, int i, int j, int k)
{
    return i >= 2 && j >= 2 && k >= 2 && i <= d.mx - 3 && j <= d.my - 3 && k <= d.mz - 3;
}

// les.c:717-758 / 1066-1107: the Cs-averaging weight of neighbour (K,J,I)
// of centre row j (the get_weight value, 0 on non-periodic ghost planes,
// where the j test reads the centre's j, so a j = my-1 neighbour keeps its
// weight) and the point its value is read from (periodic wrap)
__device__ __forceinline__ double cs_weight(const Dims &d, const double *__restrict__ wt, int j, int K, int J, int I,
                                            unsigned &src)
{
    double w = ld32(wt, IX32(d, K, J, I));
    if (d.ip) { if (I == 0) I = d.mx - 2; else if (I == d.mx - 1) I = 1; }
    else if (I == 0 || I == d.mx - 1) w = 0;
    if (d.jp) { if (J == 0) J = d.my - 2; else if (J == d.my - 1) J = 1; }
    else if (J == 0 || j == d.my - 1) w = 0;
    if (d.kp) { if (K == 0) K = d.mz - 2; else if (K == d.mz - 1) K = 1; }
    else if (K == 0 || K == d.mz - 1) w = 0;
    src = IX32(d, K, J, I);
    return w;
}

// k-omega.c:597-610 Compute_du_dxyz: D[a][b] = d u_a / d x_b
__device__ __forceinline__ void du_dxyz(const Geom &g, long long t, const double dc[3], const double de[3],
                                        const double dz[3], double D[3][3])
{
    const double ajc = g.aj[t];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            D[a][b] = (dc[a] * g.csi[3 * t + b] + de[a] * g.eta[3 * t + b] + dz[a] * g.zet[3 * t + b]) * ajc;
}

// S_ab = (D_ab + D_ba) / 2 and |S| = sqrt(2 S:S), terms in the order xx, xy,
// xz, yx, ... (les.c:199-203)
__device__ __forceinline__ double strain(const double D[3][3], double S[3][3])
{
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) S[a][b] = 0.5 * (D[a][b] + D[b][a]);
    double s = S[0][0] * S[0][0];
#pragma unroll
    for (int n = 1; n < 9; n++) s += S[n / 3][n % 3] * S[n / 3][n % 3];
    return sqrt(2.0 * s);
}

__device__ __forceinline__ bool interior(const Dims &d, int i, int j, int k)
{
    return i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2;
}

// one thread a point; 32-bit point and element indices (loads as uniform
// base + 32-bit offset), valid while 24 P < 2^32 (checked on the host)
#define DSM_PREAMBLE                                                                                 \
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;                                        \
    if (t >= (unsigned)d.P) return;                                                                  \
    const int i = (int)(t % (unsigned)d.mx), j = (int)((t / (unsigned)d.mx) % (unsigned)d.my),       \
              k = (int)(t / ((unsigned)d.mx * (unsigned)d.my))

// the six S_ab |S| planes: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
__device__ __forceinline__ int ss_plane(int a, int b)
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

// les.c:188-257 (first pass).  CL (-clark): the nine gradient components
// too, as planes gr + (3 a + b) P (the reference's Ax, Ay, Az).
template <bool IK, bool CL>
__global__ void __launch_bounds__(256) k_dsm_grad(Dims d, Geom g, const double *__restrict__ nvert,
                                                  const double *__restrict__ ucat, const double *__restrict__ uq,
                                                  const double *__restrict__ wt, double *__restrict__ ss,
                                                  double *__restrict__ uf, double *__restrict__ gr)
{
    DSM_PREAMBLE;
    double *pl[6];
#pragma unroll
    for (int q = 0; q < 6; q++) pl[q] = ss + (size_t)q * d.P;
    if (!interior(d, i, j, k) || ld32(nvert, t) > 1.1) {
#pragma unroll
        for (int q = 0; q < 6; q++) st32(pl[q], t, 0.0);
#pragma unroll
        for (int q = 0; q < 3; q++) st32(uf, 3 * t + q, 0.0);
        if (CL)
#pragma unroll
            for (int q = 0; q < 9; q++) st32(gr + (size_t)q * d.P, t, 0.0);
        return;
    }
    double dc[3], de[3], dz[3], D[3][3], S[3][3];
    du_center(d, i, j, k, ucat, nvert, dc, de, dz);
    du_dxyz(g, t, dc, de, dz, D);
    const double sabs = strain(D, S);
    if (CL)
#pragma unroll
        for (int q = 0; q < 9; q++) st32(gr + (size_t)q * d.P, t, D[q / 3][q % 3]);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++) st32(pl[ss_plane(a, b)], t, S[a][b] * sabs);
    if (IK) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double *u = uq + (size_t)q * d.P;
            st32(uf, 3 * t + q, simpson_ik([&](int R, int P) { return ld32(u, t + (unsigned)slot_off(d, R, 1, P)); }));
        }
        return;
    }
    Oct v[3], w;
    OCT_FOR
    {
        const unsigned n = t + (unsigned)slot_off(d, R, Q, P);
        const double wn = ld32(wt, n);
        oct_add(w, R, Q, P, wn);
#pragma unroll
        for (int q = 0; q < 3; q++) oct_add(v[q], R, Q, P, ld32(uq + (size_t)q * d.P, n) * wn);
    }
    const double ws = oct_total(w);
#pragma unroll
    for (int q = 0; q < 3; q++) st32(uf, 3 * t + q, oct_total(v[q]) / ws);
}

// rhs2.c:538-552 Calculate_Covariant_metrics: the inverse of [csi; eta; zet]
__device__ __forceinline__ void covariant_metrics(const double gg[3][3], double G[3][3])
{
    const double a11 = gg[0][0], a12 = gg[0][1], a13 = gg[0][2];
    const double a21 = gg[1][0], a22 = gg[1][1], a23 = gg[1][2];
    const double a31 = gg[2][0], a32 = gg[2][1], a33 = gg[2][2];
    const double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    G[0][0] = (a33 * a22 - a32 * a23) / det, G[0][1] = -(a33 * a12 - a32 * a13) / det;
    G[0][2] = (a23 * a12 - a22 * a13) / det;
    G[1][0] = -(a33 * a21 - a31 * a23) / det, G[1][1] = (a33 * a11 - a31 * a13) / det;
    G[1][2] = -(a23 * a11 - a21 * a13) / det;
    G[2][0] = (a32 * a21 - a31 * a22) / det, G[2][1] = -(a32 * a11 - a31 * a12) / det;
    G[2][2] = (a22 * a11 - a21 * a12) / det;
}

// les.c:306-687 (second pass).  A block of four waves covers 64 consecutive
// points; each wave streams a share of the sixteen box filters for all 64
// (wave b < 3: <U_b u_a> for a = 0..2, U_b = u . (csi, eta, zet)_b, and
// <S|S|> plane b; wave 3: planes 3-5 and the weight sums), so a wave's role
// is uniform and its loads run along i.  The raw octant totals meet in LDS,
// and wave 0 finishes each point: the divisions by the weight total, the
// Germano terms and the contractions.  CL (-clark): two more waves filter
// the Clark stresses T_ab of the box points (with the centre cell's length
// scales, les.c:339-340, 410-420), three planes each, and wave 0 adds the
// term - N_ab M_aq G_bq, N = 4 R - <T> (les.c:509-548, 648).
#define DSM_LM_PTS 64

// the Clark stress T_ab / 12 of point n from the gradient planes
__device__ __forceinline__ double clark_t(const double *__restrict__ gr, size_t Pz, unsigned n, int a, int b,
                                          const double dl2[3])
{
    const double *ga = gr + 3 * a * Pz, *gb = gr + 3 * b * Pz;
    return (ld32(ga, n) * ld32(gb, n) * dl2[0] + ld32(ga + Pz, n) * ld32(gb + Pz, n) * dl2[1] +
            ld32(ga + 2 * Pz, n) * ld32(gb + 2 * Pz, n) * dl2[2]) / 12.;
}

// the squared length scales of cell t (Calculate_dxdydz, les.c:339-340)
__device__ __forceinline__ void clark_lengths(const Geom &g, unsigned t, double dl2[3])
{
    double cs[3], ce[3], cz[3], lx, ly, lz;
#pragma unroll
    for (int q = 0; q < 3; q++) cs[q] = ld32(g.csi, 3 * t + q), ce[q] = ld32(g.eta, 3 * t + q), cz[q] = ld32(g.zet, 3 * t + q);
    dxdydz(ld32(g.aj, t), cs, ce, cz, lx, ly, lz);
    dl2[0] = lx * lx, dl2[1] = ly * ly, dl2[2] = lz * lz;
}

// At most three waves per SIMD: at four (128 VGPRs) the 27-point box
// streams of the resident waves miss the caches more often, and the pass
// takes 1.95 ms instead of 1.45 ms at 256x128x128.
template <bool IK, bool CL>
__global__ void __launch_bounds__(CL ? 384 : 256) __attribute__((amdgpu_waves_per_eu(1, 3))) k_dsm_lm(Dims d, Geom g, const double *__restrict__ nvert,
                                                          const double *__restrict__ uq, const double *__restrict__ Ub,
                                                          const double *__restrict__ wt, const double *__restrict__ ss,
                                                          const double *__restrict__ uf, const double *__restrict__ gr,
                                                          double *__restrict__ LM, double *__restrict__ MM)
{
    // 0-8 <U_b u_a> (3b+a), 9-14 <S|S|> planes, 15 w total, 16 sum_weight, 17-22 <T> planes (-clark)
    __shared__ double sh[CL ? 23 : 17][DSM_LM_PTS];
    const unsigned lane = threadIdx.x & 63, role = threadIdx.x >> 6;
    const unsigned t = blockIdx.x * DSM_LM_PTS + lane;
    const unsigned mx = (unsigned)d.mx, my = (unsigned)d.my;
    const int i = (int)(t % mx), j = (int)((t / mx) % my), k = (int)(t / (mx * my));
    const bool act = t < (unsigned)d.P && interior(d, i, j, k) && ld32(nvert, t) <= 1.1;
    if (act && IK) {   // -testfilter_ik: the plane filters, no weights (totals of 1)
        const size_t Pz = (size_t)d.P;
        auto at = [&](int R, int P) { return t + (unsigned)slot_off(d, R, 1, P); };
        if (role < 3) {
            const double *Upl = Ub + role * Pz, *sp = ss + role * Pz;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double *ua = uq + a * Pz;
                sh[3 * role + a][lane] =
                    simpson_ik([&](int R, int P) { return ld32(Upl, at(R, P)) * ld32(ua, at(R, P)); });
            }
            sh[9 + role][lane] = simpson_ik([&](int R, int P) { return ld32(sp, at(R, P)); });
        } else if (role == 3) {
#pragma unroll
            for (int q = 3; q < 6; q++) {
                const double *pq = ss + q * Pz;
                sh[9 + q][lane] = simpson_ik([&](int R, int P) { return ld32(pq, at(R, P)); });
            }
            sh[15][lane] = 1.0;
            sh[16][lane] = 0.0;
        } else if (CL) {   // <T_ab>: (0,0) (0,1) (0,2) on wave 4, (1,1) (1,2) (2,2) on wave 5
            double dl2[3];
            clark_lengths(g, t, dl2);
            const int a = role == 4 ? 0 : 1;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b < a) continue;
                sh[17 + ss_plane(a, b)][lane] = simpson_ik([&](int R, int P) { return clark_t(gr, Pz, at(R, P), a, b, dl2); });
            }
            if (role == 5) sh[17 + ss_plane(2, 2)][lane] = simpson_ik([&](int R, int P) { return clark_t(gr, Pz, at(R, P), 2, 2, dl2); });
        }
    } else if (act) {
        const size_t Pz = (size_t)d.P;
        if (role < 3) {   // <U_b u_a>, a = 0..2, for b = role, and the S|S| plane b
            const double *Upl = Ub + role * Pz, *u0p = uq, *u1p = uq + Pz, *u2p = uq + 2 * Pz, *sp = ss + role * Pz;
            Oct v[4];
            OCT_FOR
            {
                const unsigned n = t + (unsigned)slot_off(d, R, Q, P);
                const double wn = ld32(wt, n), U = ld32(Upl, n);
                const double u0 = ld32(u0p, n), u1 = ld32(u1p, n), u2 = ld32(u2p, n);
                oct_add(v[0], R, Q, P, U * u0 * wn);
                oct_add(v[1], R, Q, P, U * u1 * wn);
                oct_add(v[2], R, Q, P, U * u2 * wn);
                oct_add(v[3], R, Q, P, ld32(sp, n) * wn);
            }
#pragma unroll
            for (int a = 0; a < 3; a++) sh[3 * role + a][lane] = oct_total(v[a]);
            sh[9 + role][lane] = oct_total(v[3]);
        } else if (role == 3) {   // S|S| planes 3-5 and the weight sums
            const double *p3 = ss + 3 * Pz, *p4 = p3 + Pz, *p5 = p4 + Pz;
            Oct v[3], w;
            OCT_FOR
            {
                const unsigned n = t + (unsigned)slot_off(d, R, Q, P);
                const double wn = ld32(wt, n);
                oct_add(v[0], R, Q, P, ld32(p3, n) * wn);
                oct_add(v[1], R, Q, P, ld32(p4, n) * wn);
                oct_add(v[2], R, Q, P, ld32(p5, n) * wn);
                oct_add(w, R, Q, P, wn);
            }
#pragma unroll
            for (int q = 0; q < 3; q++) sh[12 + q][lane] = oct_total(v[q]);
            sh[15][lane] = oct_total(w);
            // les.c:380-393: sum of weight * coef over p (i), q (j), r (k)
            double sum_weight = 0;
#pragma unroll
            for (int pp = 0; pp < 3; pp++)
#pragma unroll
                for (int q = 0; q < 3; q++)
#pragma unroll
                    for (int r = 0; r < 3; r++) {
                        const double cf = 0.125 * (double)((1 + (r == 1)) * (1 + (q == 1)) * (1 + (pp == 1)));
                        sum_weight += ld32(wt, t + (unsigned)slot_off(d, r, q, pp)) * cf;
                    }
            sh[16][lane] = sum_weight;
        } else if (CL) {   // <T_ab>: (0,0) (0,1) (0,2) on wave 4, (1,1) (1,2) (2,2) on wave 5
            double dl2[3];
            clark_lengths(g, t, dl2);
            Oct v[3];
            if (role == 4) {
                OCT_FOR
                {
                    const unsigned n = t + (unsigned)slot_off(d, R, Q, P);
                    const double wn = ld32(wt, n);
#pragma unroll
                    for (int b = 0; b < 3; b++) oct_add(v[b], R, Q, P, clark_t(gr, Pz, n, 0, b, dl2) * wn);
                }
#pragma unroll
                for (int b = 0; b < 3; b++) sh[17 + ss_plane(0, b)][lane] = oct_total(v[b]);
            } else {
                OCT_FOR
                {
                    const unsigned n = t + (unsigned)slot_off(d, R, Q, P);
                    const double wn = ld32(wt, n);
                    oct_add(v[0], R, Q, P, clark_t(gr, Pz, n, 1, 1, dl2) * wn);
                    oct_add(v[1], R, Q, P, clark_t(gr, Pz, n, 1, 2, dl2) * wn);
                    oct_add(v[2], R, Q, P, clark_t(gr, Pz, n, 2, 2, dl2) * wn);
                }
                sh[17 + ss_plane(1, 1)][lane] = oct_total(v[0]);
                sh[17 + ss_plane(1, 2)][lane] = oct_total(v[1]);
                sh[17 + ss_plane(2, 2)][lane] = oct_total(v[2]);
            }
        }
    }
    __syncthreads();
    if (role != 0 || t >= (unsigned)d.P) return;
    if (!act) {
        st32(LM, t, 0.0);
        st32(MM, t, 0.0);
        return;
    }
    const double *met[3] = {g.csi, g.eta, g.zet};
    const double ws = sh[15][lane];
    const double filter = pow(1. / ld32(g.aj, t), 1. / 3.);
    const double test_filter = IK ? pow(5.0, 1. / 3.) * filter : pow(sh[16][lane], 1. / 3.);
    const double ufc[3] = {ld32(uf, 3 * t), ld32(uf, 3 * t + 1), ld32(uf, 3 * t + 2)};
    double L[3][3], SSh[3][3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const double Uf = ufc[0] * ld32(met[b], 3 * t) + ufc[1] * ld32(met[b], 3 * t + 1) + ufc[2] * ld32(met[b], 3 * t + 2);
#pragma unroll
        for (int a = 0; a < 3; a++) L[b][a] = sh[3 * b + a][lane] / ws - Uf * ufc[a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++) SSh[a][b] = SSh[b][a] = sh[9 + ss_plane(a, b)][lane] / ws;
    double dc[3], de[3], dz[3], D[3][3], Sh[3][3];
    du_center(d, i, j, k, uf, nvert, dc, de, dz);
    du_dxyz(g, t, dc, de, dz, D);
    const double S_hat = strain(D, Sh);
    double gg[3][3], gc[3][3], G[3][3], M[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) gg[a][b] = ld32(met[a], 3 * t + b);
    covariant_metrics(gg, gc);
#pragma unroll
    for (int a = 0; a < 3; a++) G[a][a] = gc[0][a] * gc[0][a] + gc[1][a] * gc[1][a] + gc[2][a] * gc[2][a];
    G[0][1] = G[1][0] = gc[0][1] * gc[0][0] + gc[1][1] * gc[1][0] + gc[2][1] * gc[2][0];
    G[0][2] = G[2][0] = gc[0][2] * gc[0][0] + gc[1][2] * gc[1][0] + gc[2][2] * gc[2][0];
    G[1][2] = G[2][1] = gc[0][1] * gc[0][2] + gc[1][1] * gc[1][2] + gc[2][1] * gc[2][2];
    const double tf2 = test_filter * test_filter, f2 = filter * filter;   // pow(x, 2.)
    double Mc[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Mc[a][b] = -tf2 * S_hat * Sh[a][b] + f2 * SSh[a][b];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int q = 0; q < 3; q++) M[a][q] = Mc[a][0] * gg[q][0] + Mc[a][1] * gg[q][1] + Mc[a][2] * gg[q][2];
    double N[3][3];
    if (CL) {   // N = 4 R - <T> with R from the filtered velocity's gradient (les.c:509-548)
        double dl2[3], Nc[3][3];
        clark_lengths(g, t, dl2);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const double Rab = (D[a][0] * D[b][0] * dl2[0] + D[a][1] * D[b][1] * dl2[1] + D[a][2] * D[b][2] * dl2[2]) / 12.;
                Nc[a][b] = 4.0 * Rab - sh[17 + ss_plane(a, b)][lane] / ws;
            }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int q = 0; q < 3; q++) N[a][q] = Nc[a][0] * gg[q][0] + Nc[a][1] * gg[q][1] + Nc[a][2] * gg[q][2];
    }
    double num = 0, denom = 0;
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                num += L[b][a] * M[a][q] * G[b][q];
                if (CL) num -= N[a][b] * M[a][q] * G[b][q];   // les.c:648
            }
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int n = 0; n < 3; n++)
#pragma unroll
            for (int l = 0; l < 3; l++) denom += M[n][m] * M[n][l] * G[m][l];
    st32(LM, t, num);
    st32(MM, t, denom);
}

// the clipping of les.c:943-994 at point (k,j,i) with mask value nv
__device__ __forceinline__ double clip_cs(const Dims &d, int i, int j, int k, double nv, double c, double max_cs)
{
    if (!interior(d, i, j, k) || nv > 1.1) return 0;
    if (nv > 0.1 && nv < 1.1) c = PMAX(VFS_WALL_CS, c);
    return PMIN(PMAX(c, 0), max_cs);
}

// les.c:708-789 (third pass), with the clipping of les.c:943-994 when CLIP.
// FILTER: the test filter of LM and MM (not with a homogeneous direction or
// -les 3); IK: its i-k plane form; relax: -les 3 from ti = 100 on, with the
// previous Cs read from cs (in place).
template <bool FILTER, bool IK>
__global__ void __launch_bounds__(256) k_dsm_cs(Dims d, const double *__restrict__ nvert,
                                                const double *__restrict__ wt, const double *__restrict__ LM,
                                                const double *__restrict__ MM, int relax, int clip, double max_cs,
                                                double *cs)
{
    DSM_PREAMBLE;
    const double nv = ld32(nvert, t);
    if (!interior(d, i, j, k) || nv > 1.1) {
        st32(cs, t, 0.0);
        return;
    }
    double lm, mm;
    if (!FILTER) {
        lm = ld32(LM, t);
        mm = ld32(MM, t);
    } else if (IK) {
        auto src = [&](int R, int P) { unsigned s; cs_weight(d, wt, j, k + R - 1, j, i + P - 1, s); return s; };
        lm = simpson_ik([&](int R, int P) { return ld32(LM, src(R, P)); });
        mm = simpson_ik([&](int R, int P) { return ld32(MM, src(R, P)); });
    } else {
        Oct l, m, w;
        if (away_from_ghosts(d, i, j, k)) {
            OCT_FOR
            {
                const unsigned n = t + (unsigned)slot_off(d, R, Q, P);
                const double wn = ld32(wt, n);
                oct_add(w, R, Q, P, wn);
                oct_add(l, R, Q, P, ld32(LM, n) * wn);
                oct_add(m, R, Q, P, ld32(MM, n) * wn);
            }
        } else {
            OCT_FOR_RT
            {
                unsigned src;
                const double wn = cs_weight(d, wt, j, k + R - 1, j + Q - 1, i + P - 1, src);
                oct_add(w, R, Q, P, wn);
                oct_add(l, R, Q, P, ld32(LM, src) * wn);
                oct_add(m, R, Q, P, ld32(MM, src) * wn);
            }
        }
        const double ws = oct_total(w);
        lm = oct_total(l) / ws;
        mm = oct_total(m) / ws;
    }
    double C = 0.5 * lm / (mm + VFS_LES_EPS);
    if (relax) C = (1.0 - 0.001) * ld32(cs, t) + 0.001 * C;   // les.c:777-782
    const double c = PMAX(C, 0);
    st32(cs, t, clip ? clip_cs(d, i, j, k, nv, c, max_cs) : c);
}

// les.c:791-942: sums of LM and MM over the fluid cells of each i-k plane
// (dir 3; one thread a j) or each line along direction dir (one thread a
// line (p1, p2): (k,j) for i, (k,i) for j, (j,i) for k), in the reference's
// order; sums[3 n + pos]: LM, MM and the count of line pos of n
__global__ void k_homo_sums(Dims d, int dir, const double *__restrict__ nvert, const double *__restrict__ LM,
                            const double *__restrict__ MM, double *__restrict__ sums)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sl = 0, sm = 0, cnt = 0;
    if (dir == 3) {
        if (t >= d.my) return;
        const int j = (int)t;
        if (j >= 1 && j <= d.my - 2)
            for (int k = 1; k < d.mz - 1; k++)
                for (int i = 1; i < d.mx - 1; i++) {
                    const long long id = IX(d, k, j, i);
                    if (nvert[id] < 0.1) { sl += LM[id]; sm += MM[id]; cnt += 1; }
                }
        sums[t] = sl; sums[d.my + t] = sm; sums[2 * d.my + t] = cnt;
        return;
    }
    const int n1 = dir == 2 ? d.my : d.mz, n2 = dir == 0 ? d.my : d.mx, len = dir == 0 ? d.mx : (dir == 1 ? d.my : d.mz);
    const long long nl = (long long)n1 * n2;
    if (t >= nl) return;
    const int p1 = (int)(t / n2), p2 = (int)(t % n2);
    if (p1 >= 1 && p1 <= n1 - 2 && p2 >= 1 && p2 <= n2 - 2)
        for (int q = 1; q < len - 1; q++) {
            const int k = dir == 2 ? q : p1, j = dir == 0 ? p2 : (dir == 1 ? q : p1), i = dir == 0 ? q : p2;
            const long long id = IX(d, k, j, i);
            if (nvert[id] < 0.1) { sl += LM[id]; sm += MM[id]; cnt += 1; }
        }
    sums[t] = sl; sums[nl + t] = sm; sums[2 * nl + t] = cnt;
}

// the averages (sums / count where count > 0) into Cs: every interior cell
// for the i-k planes, the fluid cells for lines; then the clipping of every
// point (les.c:943-994)
__global__ void k_homo_assign(Dims d, int dir, const double *__restrict__ nvert, const double *__restrict__ sums,
                              double max_cs, double *__restrict__ cs)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    const double nv = nvert[t];
    double c = cs[t];
    if (interior(d, i, j, k)) {
        long long pos, nl;
        if (dir == 3) { pos = j; nl = d.my; }
        else {
            const int n2 = dir == 0 ? d.my : d.mx;
            const int p1 = dir == 2 ? j : k, p2 = dir == 0 ? j : i;
            pos = (long long)p1 * n2 + p2;
            nl = (long long)(dir == 2 ? d.my : d.mz) * n2;
        }
        if (dir == 3 || nv < 0.1) {
            double a = sums[pos], b = sums[nl + pos];
            const double n = sums[2 * nl + pos];
            if (n > 0) { a /= n; b /= n; }
            c = 0.5 * a / (b + VFS_LES_EPS);
        }
    }
    cs[t] = clip_cs(d, i, j, k, nv, c, max_cs);
}

// les.c:1049-1107: the box filter of the clipped Cs at the fluid cells
template <bool IK>
__global__ void __launch_bounds__(256) k_dsm_smooth(Dims d, const double *__restrict__ nvert, const double *__restrict__ wt,
                             const double *__restrict__ cst, double *__restrict__ cs)
{
    DSM_PREAMBLE;
    if (!interior(d, i, j, k) || !(ld32(nvert, t) < 0.1)) {
        st32(cs, t, ld32(cst, t));
        return;
    }
    if (IK) {
        auto src = [&](int R, int P) { unsigned s; cs_weight(d, wt, j, k + R - 1, j, i + P - 1, s); return s; };
        st32(cs, t, simpson_ik([&](int R, int P) { return ld32(cst, src(R, P)); }));
        return;
    }
    Oct v, w;
    OCT_FOR
    {
        unsigned src;
        const double wn = cs_weight(d, wt, j, k + R - 1, j + Q - 1, i + P - 1, src);
        oct_add(w, R, Q, P, wn);
        oct_add(v, R, Q, P, ld32(cst, src) * wn);
    }
    st32(cs, t, oct_total(v) / oct_total(w));
}

__global__ void k_fill(long long n, double v, double *__restrict__ x)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = v;
}

// periodic ghost copies of np consecutive scalar planes of P points
__global__ void k_periodic_planes(Dims d, int np, double *v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i, a, b, c;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i) || !per_src(d, i, j, k, a, b, c)) return;
    const long long o = IX(d, k, j, i), s = IX(d, c, b, a);
    for (int p = 0; p < np; p++) v[p * d.P + o] = v[p * d.P + s];
}

static int periodic_planes(vfs_ctx *c, double *v, int np)
{
    if (!(c->d.ip || c->d.jp || c->d.kp)) return 0;
    const long long n = shell_count(c->d);
    hipLaunchKernelGGL(k_periodic_planes, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->d, np, v);
    VFS_CHECK_LAUNCH();
    return 0;
}

extern "C" int vfs_smagorinsky_constant(vfs_ctx *c, const double *ucat, const vfs_les_options *o, int ti, int tistart,
                                        int rstart, double *cs)
{
    VFS_REQUIRE(o, "options required");
    VFS_REQUIRE(o->les >= 1 && o->les <= 3, "les must be 1 (constant Cs), 2 (dynamic) or 3 (dynamic, relaxed)");
    VFS_REQUIRE(ucat && cs, "ucat and cs are required");
    const int nh = (o->i_homo_filter != 0) + (o->j_homo_filter != 0) + (o->k_homo_filter != 0);
    VFS_REQUIRE(nh <= 1 || (nh == 2 && o->i_homo_filter && o->k_homo_filter),
                "homogeneous filters: one direction, or i and k together");
    const Dims d = c->d;
    VFS_REQUIRE(VFS_IX32_OK(d.P), "grid too large for the dynamic procedure's 32-bit indexing (24 P >= 2^32)");
    const bool fresh = ti < 2 && tistart == 0 && !rstart;
    if (fresh || o->les == 1) {   // les.c:69-77
        hipLaunchKernelGGL(k_fill, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d.P, fresh ? 0.0 : 0.01, cs);
        VFS_CHECK_LAUNCH();
        return 0;
    }
    // 19 planes, and the 9 gradient planes of -clark
    const bool cl = d.clark != 0;
    if (!c->dsm) VFS_CHECK_HIP(hipMalloc(&c->dsm, (size_t)(cl ? 28 : 19) * d.P * sizeof(double)));
    const size_t P = (size_t)d.P;
    double *ss = c->dsm, *uf = ss + 6 * P, *LM = uf + 3 * P, *MM = LM + P, *cst = MM + P, *wt = cst + P;
    double *uq = wt + P, *Ub = uq + 3 * P, *gr = cl ? Ub + 3 * P : nullptr;
    const bool ik = o->testfilter_ik != 0, les3 = o->les == 3, homo = nh > 0;
    const dim3 blk(256), grd(grid_for(d.P, 256));
    hipLaunchKernelGGL(k_dsm_point, grd, blk, 0, c->stream, d, c->g, c->nvert, ucat, wt, uq, Ub);
    VFS_CHECK_LAUNCH();
#define DSM_GRAD(IKV, CLV) \
    hipLaunchKernelGGL((k_dsm_grad<IKV, CLV>), grd, blk, 0, c->stream, d, c->g, c->nvert, ucat, uq, wt, ss, uf, gr)
    if (ik) { if (cl) DSM_GRAD(true, true); else DSM_GRAD(true, false); }
    else { if (cl) DSM_GRAD(false, true); else DSM_GRAD(false, false); }
#undef DSM_GRAD
    VFS_CHECK_LAUNCH();
    if (periodic_planes(c, ss, 6) || vfs_periodic_vec(c, uf)) return -1;
    if (cl && periodic_planes(c, gr, 9)) return -1;
    const dim3 glm(grid_for(d.P, DSM_LM_PTS));
#define DSM_LM(IKV, CLV) \
    hipLaunchKernelGGL((k_dsm_lm<IKV, CLV>), glm, dim3(CLV ? 384 : 256), 0, c->stream, d, c->g, c->nvert, uq, Ub, wt, \
                       ss, uf, gr, LM, MM)
    if (ik) { if (cl) DSM_LM(true, true); else DSM_LM(true, false); }
    else { if (cl) DSM_LM(false, true); else DSM_LM(false, false); }
#undef DSM_LM
    VFS_CHECK_LAUNCH();
    if (periodic_planes(c, LM, 2)) return -1;   // LM and MM are adjacent planes
    const double max_cs = o->max_cs;
    if (homo || les3) {
        // LM and MM unfiltered; -les 3 relaxes in place and skips the averages
        const int relax = les3 && !(ti < 100 && tistart == 0 && !rstart);
        hipLaunchKernelGGL((k_dsm_cs<false, false>), grd, blk, 0, c->stream, d, c->nvert, wt, LM, MM, relax,
                           les3 ? 1 : 0, max_cs, cs);
        VFS_CHECK_LAUNCH();
        if (!les3) {
            const int dir = nh == 2 ? 3 : (o->i_homo_filter ? 0 : (o->j_homo_filter ? 1 : 2));
            const long long nl = dir == 3 ? d.my : (long long)(dir == 2 ? d.my : d.mz) * (dir == 0 ? d.my : d.mx);
            double *sums = cst;   // 3 nl doubles: cst and wt (2 P) have room
            VFS_REQUIRE(3 * nl <= 2 * (long long)P, "grid too small for the line sums");
            hipLaunchKernelGGL(k_homo_sums, dim3(grid_for(nl, 64)), dim3(64), 0, c->stream, d, dir, c->nvert, LM, MM,
                               sums);
            VFS_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_homo_assign, grd, blk, 0, c->stream, d, dir, c->nvert, sums, max_cs, cs);
            VFS_CHECK_LAUNCH();
        }
        return vfs_periodic_sca(c, cs);
    }
    if (ik) hipLaunchKernelGGL((k_dsm_cs<true, true>), grd, blk, 0, c->stream, d, c->nvert, wt, LM, MM, 0, 1, max_cs, cst);
    else hipLaunchKernelGGL((k_dsm_cs<true, false>), grd, blk, 0, c->stream, d, c->nvert, wt, LM, MM, 0, 1, max_cs, cst);
    VFS_CHECK_LAUNCH();
    if (vfs_periodic_sca(c, cst)) return -1;
    if (ik) hipLaunchKernelGGL(k_dsm_smooth<true>, grd, blk, 0, c->stream, d, c->nvert, wt, cst, cs);
    else hipLaunchKernelGGL(k_dsm_smooth<false>, grd, blk, 0, c->stream, d, c->nvert, wt, cst, cs);
    VFS_CHECK_LAUNCH();
    return 0;
}
