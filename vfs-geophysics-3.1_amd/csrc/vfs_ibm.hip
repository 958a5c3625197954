// CURVIB velocity reconstruction at immersed-boundary nodes.
//
// vfs_ib_search builds the node list of ibm_search_advanced (ibm.c:366-497)
// for a given mask: for every IB node the nearest point of a triangulated
// body (nearestcell1, ibm.c:1366-1511) and where the ray from that point
// through the node meets the next patch of cell centres (InterceptionPoint,
// ibm.c:1759-2250).  vfs_ib_interpolate then runs the per-step part,
// ibm_interpolation_advanced (ibm.c:3294-3698): the velocity and pressure at
// each node from the values interpolated at the interception point, the
// fluxes of the IB faces, and Contra2Cart.
//
// One thread per node runs the reference's element scan in element order, so
// the bounding-sphere cut and the strict-minimum updates see the same dmin
// sequence as the reference and the nearest element is the same one.  The
// per-step pass reads the old fields and writes the new node values through
// a compact staging array, as the reference reads lUcat / lP while it writes
// Ucat / P.
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "vfs_internal.h"

namespace {

struct IbSurf {
    int n_v, n_elmt;
    const double *x, *y, *z;
    const int *nv1, *nv2, *nv3;
    const double *nfx, *nfy, *nfz;
    const double *u;
};

IbSurf to_surf(const vfs_ib_surface *s)
{
    return IbSurf{s->n_v, s->n_elmt, s->x_bp, s->y_bp, s->z_bp, s->nv1, s->nv2, s->nv3, s->nf_x, s->nf_y, s->nf_z, s->u};
}

// ibm.c:28 Dist
__device__ __forceinline__ double dist3(const double a[3], const double b[3])
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

__device__ __forceinline__ void vertex(const IbSurf &s, int v, double p[3])
{
    p[0] = s.x[v];
    p[1] = s.y[v];
    p[2] = s.z[v];
}

// ibm.c:30-33 Cross
__device__ __forceinline__ void cross3(double r[3], const double a[3], const double b[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// ibm.c:1276-1364 BoundingSphere: the circumcentre of an acute element, the
// middle of the longest edge otherwise
__global__ void k_ib_bound(IbSurf s, double *__restrict__ qv, double *__restrict__ rad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= s.n_elmt) return;
    double p1[3], p2[3], p3[3], pa[3], pc[3], pf[3], pu[3], pv[3], pt[3], pd[3];
    const double p0[3] = {0, 0, 0};
    vertex(s, s.nv1[e], p1);
    vertex(s, s.nv2[e], p2);
    vertex(s, s.nv3[e], p3);
    const double l12 = dist3(p1, p2), l23 = dist3(p2, p3), l31 = dist3(p3, p1);
    const double *a, *b, *c;
    if (l12 > l23) {
        if (l12 > l31) a = p1, b = p2, c = p3;
        else a = p3, b = p1, c = p2;
    } else {
        if (l31 < l23) a = p2, b = p3, c = p1;
        else a = p3, b = p1, c = p2;
    }
    for (int q = 0; q < 3; q++) {
        pa[q] = a[q];
        pc[q] = c[q];
        pf[q] = 0.5 * (a[q] + b[q]);
    }
    for (int q = 0; q < 3; q++) pu[q] = pa[q] - pf[q], pv[q] = pc[q] - pf[q];
    cross3(pt, pu, pv);
    cross3(pd, pt, pu);
    double gama = -(dist3(pu, p0) * dist3(pu, p0) - dist3(pv, p0) * dist3(pv, p0));
    for (int q = 0; q < 3; q++) pt[q] = pv[q] - pu[q];
    double lamda = 2 * (pd[0] * pt[0] + pd[1] * pt[1] + pd[2] * pt[2]);
    gama /= lamda;
    lamda = gama < 0 ? 0 : gama;
    double qq[3];
    for (int q = 0; q < 3; q++) qq[q] = pf[q] + lamda * pd[q];
    qv[3 * e] = qq[0];
    qv[3 * e + 1] = qq[1];
    qv[3 * e + 2] = qq[2];
    rad[e] = dist3(qq, pa);
}

// compgeom.c:56-110 ISSameSide2D as it runs: the t1 / t2 of its last block
// decide, a point within |t1| / |p1p2| < 1e-10 of the line counts as inside
__device__ __forceinline__ int same_side(const double p[2], const double p1[2], const double p2[2],
                                         const double p3[2])
{
    const double t1 = (p[0] - p1[0]) * (p2[1] - p1[1]) - (p[1] - p1[1]) * (p2[0] - p1[0]);
    const double t2 = (p3[0] - p1[0]) * (p2[1] - p1[1]) - (p3[1] - p1[1]) * (p2[0] - p1[0]);
    const double lt = sqrt((p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]));
    if (fabs(t1 / lt) < 1.e-10) return 1;
    if (t1 > 0) return t2 > 0 ? 1 : -1;
    return t2 < 0 ? 1 : -1;
}

// compgeom.c:112-158 ISPointInTriangle: 1 inside or on an edge, -1 outside,
// on the coordinate plane the normal picks
__device__ int point_in_triangle(const double p[3], const double p1[3], const double p2[3], const double p3[3],
                                 double nx, double ny, double nz)
{
    int u, v;
    if (fabs(nz) >= fabs(nx) && fabs(nz) >= fabs(ny)) u = 0, v = 1;
    else if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz)) u = 2, v = 1;
    else u = 0, v = 2;
    const double a[2] = {p[u], p[v]}, b1[2] = {p1[u], p1[v]}, b2[2] = {p2[u], p2[v]}, b3[2] = {p3[u], p3[v]};
    int ls = same_side(a, b1, b2, b3);
    if (ls < 0) return ls;
    ls = same_side(a, b2, b3, b1);
    if (ls < 0) return ls;
    return same_side(a, b3, b1, b2);
}

// compgeom.c:160-186 Dis_P_Line
__device__ void dis_p_line(const double p[3], const double p1[3], const double p2[3], double po[3], double &d)
{
    const double dx21 = p2[0] - p1[0], dy21 = p2[1] - p1[1], dz21 = p2[2] - p1[2];
    const double dx31 = p[0] - p1[0], dy31 = p[1] - p1[1], dz31 = p[2] - p1[2];
    const double t = (dx31 * dx21 + dy31 * dy21 + dz31 * dz21) / (dx21 * dx21 + dy21 * dy21 + dz21 * dz21);
    if (t < 0) {
        po[0] = p1[0], po[1] = p1[1], po[2] = p1[2];
        d = sqrt(dx31 * dx31 + dy31 * dy31 + dz31 * dz31);
        return;
    }
    if (t > 1) po[0] = p2[0], po[1] = p2[1], po[2] = p2[2];
    else po[0] = p1[0] + t * dx21, po[1] = p1[1] + t * dy21, po[2] = p1[2] + t * dz21;
    d = sqrt((p[0] - po[0]) * (p[0] - po[0]) + (p[1] - po[1]) * (p[1] - po[1]) + (p[2] - po[2]) * (p[2] - po[2]));
}

// compgeom.c:188-230 triangle_intp2 / triangle_intp: barycentric weights on
// the plane the normal picks (x-dominant (y, z), y-dominant (x, z),
// z-dominant (y, x)); a NaN normal leaves w as it was
__device__ void triangle_intp(const double p[3], const double p1[3], const double p2[3], const double p3[3],
                              double nx, double ny, double nz, double w[3])
{
    int u, v;
    if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz)) u = 1, v = 2;
    else if (fabs(ny) >= fabs(nx) && fabs(ny) >= fabs(nz)) u = 0, v = 2;
    else if (fabs(nz) >= fabs(ny) && fabs(nz) >= fabs(nx)) u = 1, v = 0;
    else return;
    const double x13 = p1[u] - p3[u], y13 = p1[v] - p3[v];
    const double x23 = p2[u] - p3[u], y23 = p2[v] - p3[v];
    const double xp3 = p[u] - p3[u], yp3 = p[v] - p3[v];
    const double a = x13 * y23 - x23 * y13;
    w[0] = (y23 * xp3 - x23 * yp3) / a;
    w[1] = (-y13 * xp3 + x13 * yp3) / a;
    w[2] = 1. - w[0] - w[1];
}

// ibm.c:1670-1677: the eight triangles of a 3x3 patch; the interpolation
// cells are the vertices of the triangle hit (the switch blocks of
// InterceptionPoint list the same triples)
__constant__ int c_tri[8][3] = {{0, 1, 4}, {1, 2, 4}, {2, 4, 5}, {4, 5, 8}, {4, 7, 8}, {4, 6, 7}, {3, 4, 6}, {3, 4, 0}};

// cell of point n of patch pl (i-1, i+1, j-1, j+1, k-1, k+1 planes in turn)
__device__ __forceinline__ void patch_point(int pl, int n, int i, int j, int k, int &ii, int &jj, int &kk)
{
    const int a = n % 3 - 1, b = n / 3 - 1, s = (pl & 1) ? 1 : -1;
    if (pl < 2) ii = i + s, jj = j + a, kk = k + b;
    else if (pl < 4) ii = i + a, jj = j + s, kk = k + b;
    else ii = i + a, jj = j + b, kk = k + s;
}

// the mask of the body's nodes: nvert == 1 off the wall-function planes,
// which ibm.c:518-533 sets only after the list is made
__device__ __forceinline__ int is_node(const Dims &d, const double *nvert, long long t)
{
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || i > d.mx - 2 || j < 1 || j > d.my - 2 || k < 1 || k > d.mz - 2) return 0;
    if ((int)(nvert[t] + 0.5) != 1) return 0;
    const int *bt = d.bc;
    if (((bt[0] == -1 || bt[0] == -2) && i == 1) || ((bt[1] == -1 || bt[1] == -2) && i == d.mx - 2)) return 0;
    if (((bt[2] == -1 || bt[2] == -2) && j == 1) || ((bt[3] == -1 || bt[3] == -2) && j == d.my - 2)) return 0;
    if (((bt[4] == -1 || bt[4] == -2) && k == 1) || ((bt[5] == -1 || bt[5] == -2) && k == d.mz - 2)) return 0;
    return 1;
}

constexpr int LIST_BLOCK = 256;

__global__ void __launch_bounds__(LIST_BLOCK) k_ib_count(Dims d, const double *__restrict__ nvert, int *__restrict__ cnt)
{
    const long long t = (long long)blockIdx.x * LIST_BLOCK + threadIdx.x;
    const int f = t < d.P ? is_node(d, nvert, t) : 0;
    const int c = __syncthreads_count(f);
    if (threadIdx.x == 0) cnt[blockIdx.x] = c;
}

// node at linear position r (k, j, i order) goes to list entry n-1-r
__global__ void __launch_bounds__(LIST_BLOCK) k_ib_place(Dims d, const double *__restrict__ nvert,
                                                        const int *__restrict__ off, int n,
                                                        vfs_ib_node *__restrict__ nodes)
{
    __shared__ int s[LIST_BLOCK];
    const long long t = (long long)blockIdx.x * LIST_BLOCK + threadIdx.x;
    const int f = t < d.P ? is_node(d, nvert, t) : 0;
    s[threadIdx.x] = f;
    __syncthreads();
    for (int w = 1; w < LIST_BLOCK; w <<= 1) {   // inclusive scan
        const int v = threadIdx.x >= (unsigned)w ? s[threadIdx.x - w] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (!f) return;
    const int r = off[blockIdx.x] + s[threadIdx.x] - 1;
    vfs_ib_node nd = {};
    nd.i = (int)(t % d.mx);
    nd.j = (int)((t / d.mx) % d.my);
    nd.k = (int)(t / ((long long)d.mx * d.my));
    nodes[n - 1 - r] = nd;
}

// nearestcell1 (ibm.c:1366-1511), InterceptionPoint (ibm.c:1759-2250) with
// ICP (ibm.c:1659-1757), and the boundary fallback (ibm.c:413-493).
// One wavefront a node.  The reference scans the elements in order, skips
// one whose bounding sphere is no nearer than the running minimum, and
// takes a candidate only when it is strictly nearer (the first of equal
// candidates wins).  Here the lanes take 64 elements at a time: each one
// evaluates its element against the minimum at the start of the batch (a
// larger bound, so no element the scan would evaluate is left out) and
// keeps its nearest candidate, the first of equal ones.  The elements that
// could still win are then replayed in element order with the scan's two
// tests, so the node gets the element, point and distance of the scan.
__global__ void __launch_bounds__(64) k_ib_search(Dims d, Geom g, IbSurf s, const double *__restrict__ qv,
                                                  const double *__restrict__ rad, int n,
                                                  vfs_ib_node *__restrict__ nodes, int *__restrict__ fail)
{
    const int l = blockIdx.x, lane = threadIdx.x;
    if (l >= n) return;
    vfs_ib_node nd = nodes[l];
    const int i = nd.i, j = nd.j, k = nd.k;
    const long long id = IX(d, k, j, i);
    const double p[3] = {g.cent[3 * id], g.cent[3 * id + 1], g.cent[3 * id + 2]};

    double dmin = 1.e20, pmin[3] = {0, 0, 0};   // the same in every lane
    int cell_min = -100;
    for (int e0 = 0; e0 < s.n_elmt; e0 += 64) {
        const int e = e0 + lane;
        double lb = 0, v = INFINITY, po[3] = {0, 0, 0};
        bool cand = false;
        if (e < s.n_elmt) {
            const double q[3] = {qv[3 * e], qv[3 * e + 1], qv[3 * e + 2]};
            lb = dist3(p, q) - rad[e];
            if (lb < dmin) {
                const double nfx = s.nfx[e], nfy = s.nfy[e], nfz = s.nfz[e];
                double p1[3], p2[3], p3[3];
                vertex(s, s.nv1[e], p1);
                vertex(s, s.nv2[e], p2);
                vertex(s, s.nv3[e], p3);
                double tf = ((p[0] - p1[0]) * nfx + (p[1] - p1[1]) * nfy + (p[2] - p1[2]) * nfz);
                if (fabs(tf) < 1.e-10) tf = 1.e-10;
                if (tf >= 0) {
                    const double pj[3] = {p[0] - tf * nfx, p[1] - tf * nfy, p[2] - tf * nfz};
                    if (point_in_triangle(pj, p1, p2, p3, nfx, nfy, nfz) == 1) {
                        v = tf, po[0] = pj[0], po[1] = pj[1], po[2] = pj[2];
                    } else {
                        double pe[3], dd;
                        dis_p_line(p, p1, p2, pe, dd);
                        if (dd < v) v = dd, po[0] = pe[0], po[1] = pe[1], po[2] = pe[2];
                        dis_p_line(p, p2, p3, pe, dd);
                        if (dd < v) v = dd, po[0] = pe[0], po[1] = pe[1], po[2] = pe[2];
                        dis_p_line(p, p3, p1, pe, dd);
                        if (dd < v) v = dd, po[0] = pe[0], po[1] = pe[1], po[2] = pe[2];
                    }
                    cand = v < dmin;
                }
            }
        }
        unsigned long long bal = __ballot(cand);
        while (bal) {
            const int src = __ffsll(bal) - 1;   // the earliest element left
            bal &= bal - 1;
            const double lbs = __shfl(lb, src), vs = __shfl(v, src);
            const double x0 = __shfl(po[0], src), x1 = __shfl(po[1], src), x2 = __shfl(po[2], src);
            if (lbs < dmin && vs < dmin) dmin = vs, pmin[0] = x0, pmin[1] = x1, pmin[2] = x2, cell_min = e0 + src;
        }
    }
    if (cell_min == -100) {
        if (lane == 0) {
            atomicAdd(fail, 1);
            nodes[l] = nd;
        }
        return;
    }
    nd.cell = cell_min;
    nd.pmin[0] = pmin[0], nd.pmin[1] = pmin[1], nd.pmin[2] = pmin[2];
    nd.d_s = dmin;
    {
        double p1[3], p2[3], p3[3], w[3] = {nd.cs1, nd.cs2, nd.cs3};
        vertex(s, s.nv1[cell_min], p1);
        vertex(s, s.nv2[cell_min], p2);
        vertex(s, s.nv3[cell_min], p3);
        triangle_intp(pmin, p1, p2, p3, s.nfx[cell_min], s.nfy[cell_min], s.nfz[cell_min], w);
        nd.cs1 = w[0], nd.cs2 = w[1], nd.cs3 = w[2];
    }

    double nr[3] = {p[0] - pmin[0], p[1] - pmin[1], p[2] - pmin[2]};
    const double dr = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    nr[0] /= dr, nr[1] /= dr, nr[2] /= dr;
    nd.imode = -100;
    for (int pl = 0; pl < 6 && nd.imode < 0; pl++) {
        double pc[9][3];
        for (int q = 0; q < 9; q++) {
            int ii, jj, kk;
            patch_point(pl, q, i, j, k, ii, jj, kk);
            const long long c = IX(d, kk, jj, ii);
            pc[q][0] = g.cent[3 * c], pc[q][1] = g.cent[3 * c + 1], pc[q][2] = g.cent[3 * c + 2];
        }
        for (int t = 0; t < 8; t++) {
            const double *p1 = pc[c_tri[t][0]], *p2 = pc[c_tri[t][1]], *p3 = pc[c_tri[t][2]];
            const double dx1 = p[0] - p1[0], dy1 = p[1] - p1[1], dz1 = p[2] - p1[2];
            const double dx2 = p2[0] - p1[0], dy2 = p2[1] - p1[1], dz2 = p2[2] - p1[2];
            const double dx3 = p3[0] - p1[0], dy3 = p3[1] - p1[1], dz3 = p3[2] - p1[2];
            double dd = (nr[0] * (dy2 * dz3 - dz2 * dy3) - nr[1] * (dx2 * dz3 - dz2 * dx3) +
                         nr[2] * (dx2 * dy3 - dy2 * dx3));
            if (!(fabs(dd) > 1.e-10)) continue;
            dd = -(dx1 * (dy2 * dz3 - dz2 * dy3) - dy1 * (dx2 * dz3 - dz2 * dx3) + dz1 * (dx2 * dy3 - dy2 * dx3)) / dd;
            if (!(dd > 0)) continue;
            const double pint[3] = {p[0] + dd * nr[0], p[1] + dd * nr[1], p[2] + dd * nr[2]};
            const double rx1 = p2[0] - p1[0], ry1 = p2[1] - p1[1], rz1 = p2[2] - p1[2];
            const double rx2 = p3[0] - p1[0], ry2 = p3[1] - p1[1], rz2 = p3[2] - p1[2];
            const double nxt = ry1 * rz2 - rz1 * ry2, nyt = -rx1 * rz2 + rz1 * rx2, nzt = rx1 * ry2 - ry1 * rx2;
            if (point_in_triangle(pint, p1, p2, p3, nxt, nyt, nzt) < 0) continue;
            double w[3] = {nd.cr1, nd.cr2, nd.cr3};
            triangle_intp(pint, p1, p2, p3, nxt, nyt, nzt, w);
            nd.cr1 = w[0], nd.cr2 = w[1], nd.cr3 = w[2];
            nd.d_i = sqrt((pint[0] - p[0]) * (pint[0] - p[0]) + (pint[1] - p[1]) * (pint[1] - p[1]) +
                          (pint[2] - p[2]) * (pint[2] - p[2]));
            nd.imode = t;
            int ii, jj, kk;
            patch_point(pl, c_tri[t][0], i, j, k, ii, jj, kk);
            nd.i1 = ii, nd.j1 = jj, nd.k1 = kk;
            patch_point(pl, c_tri[t][1], i, j, k, ii, jj, kk);
            nd.i2 = ii, nd.j2 = jj, nd.k2 = kk;
            patch_point(pl, c_tri[t][2], i, j, k, ii, jj, kk);
            nd.i3 = ii, nd.j3 = jj, nd.k3 = kk;
            break;
        }
    }
    if (nd.imode < 0) {
        int ti, tj, tk;
        if (i == 1 || i == d.mx - 2 || j == 1 || j == d.my - 2) ti = i, tj = j, tk = s.nfz[cell_min] > 0 ? k + 1 : k - 1;
        else if (k == 1 || k == d.mz - 2) ti = i, tj = j + 1, tk = k;
        else {
            if (lane == 0) {
                atomicAdd(fail, 1);
                nodes[l] = nd;
            }
            return;
        }
        const long long c = IX(d, tk, tj, ti);
        const double pt[3] = {g.cent[3 * c], g.cent[3 * c + 1], g.cent[3 * c + 2]};
        nd.d_i = sqrt((p[0] - pt[0]) * (p[0] - pt[0]) + (p[1] - pt[1]) * (p[1] - pt[1]) + (p[2] - pt[2]) * (p[2] - pt[2]));
        nd.cr1 = 1., nd.cr2 = 0., nd.cr3 = 0.;
        nd.i1 = nd.i2 = nd.i3 = ti;
        nd.j1 = nd.j2 = nd.j3 = tj;
        nd.k1 = nd.k2 = nd.k3 = tk;
    }
    if (lane == 0) nodes[l] = nd;
}

// sediment_transport.c:18-68: the rough-wall law of the IB wall model
__device__ double hy_E(double utau, double ks, double nu)
{
    const double kplus = utau * ks / nu, kappa = 0.41, B = 5.5;
    double dB;
    if (kplus <= 2.25) dB = 0.0;
    else if (kplus > 2.25 && kplus < 90.) dB = (B - 8.5 + 1. / kappa * log(kplus)) * sin(0.4258 * (log(kplus) - 0.811));
    else dB = B - 8.5 + 1. / kappa * log(kplus);
    return exp(kappa * (B - dB));
}

__device__ double hy_u(double nu, double y, double utau, double ks)
{
    const double y0plus = 11.53, kappa = 0.41, yplus = utau * y / nu;
    if (yplus <= y0plus) return utau * yplus;
    return utau / kappa * log(hy_E(utau, ks, nu) * yplus);
}

__device__ double hy_f(double nu, double u, double y, double utau0, double ks)
{
    const double y0plus = 11.53, kappa = 0.41, yplus = utau0 * y / nu;
    if (yplus <= y0plus) return utau0 * yplus - u;
    return utau0 * (1. / kappa * log(hy_E(utau0, ks, nu) * yplus)) - u;
}

// find_utau_hydset: Newton with a central-difference slope, 30 steps at most
__device__ double hy_utau(double nu, double u, double y, double guess, double ks)
{
    double utau = guess, utau0 = guess;
    for (int ii = 0; ii < 30; ii++) {
        const double eps = 1.e-7;
        const double df = (hy_f(nu, u, y, utau0 + eps, ks) - hy_f(nu, u, y, utau0 - eps, ks)) / (2. * eps);
        utau = utau0 - hy_f(nu, u, y, utau0, ks) / df;
        if (fabs(utau - utau0) < 1.e-7) break;
        utau0 = utau;
    }
    return utau;
}

// u_loglaw_roughness (wallfunction.c:221-231)
__device__ double rl_u(double y, double utau, double ks)
{
    const double kappa = 0.41;
    return utau * (1. / kappa * log(y / ks) + 8.5);
}

// find_utau_loglaw_roughness (wallfunction.c:423-437): Newton on
// f = rl_u - u with a central-difference slope, 30 steps at most
__device__ double rl_utau(double u, double y, double guess, double ks)
{
    double x = guess, x0 = guess;
    for (int ii = 0; ii < 30; ii++) {
        const double eps = 1.e-7;
        const double df = ((rl_u(y, x0 + eps, ks) - u) - (rl_u(y, x0 - eps, ks) - u)) / (2 * eps);
        x = x0 - (rl_u(y, x0, ks) - u) / df;
        if (fabs(x0 - x) < 1.e-10) break;
        x0 = x;
    }
    return x;
}

// a node record that fits the grid and the body (entries that do not are
// skipped, so no kernel reads outside its arrays)
__device__ __forceinline__ bool node_ok(const Dims &d, int n_elmt, const vfs_ib_node &nd)
{
    auto in = [&](int i, int j, int k, int lo) {
        return i >= lo && i <= d.mx - 1 - lo && j >= lo && j <= d.my - 1 - lo && k >= lo && k <= d.mz - 1 - lo;
    };
    return nd.cell >= 0 && nd.cell < n_elmt && in(nd.i, nd.j, nd.k, 1) && in(nd.i1, nd.j1, nd.k1, 0) &&
           in(nd.i2, nd.j2, nd.k2, 0) && in(nd.i3, nd.j3, nd.k3, 0);
}

// ibm.c:3390-3530: the node values from the old fields into the staging
// array (u x|y|z, p in blocks of n); ustar of the node in the field
__global__ void __launch_bounds__(256) k_ib_nodes(Dims d, const double *__restrict__ nvert, IbSurf s,
                                                  const vfs_ib_node *__restrict__ nodes, int n,
                                                  const double *__restrict__ ucat, const double *__restrict__ pf,
                                                  vfs_ib_options opt, double *__restrict__ ustar,
                                                  double *__restrict__ stage)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const vfs_ib_node nd = nodes[l];
    if (!node_ok(d, s.n_elmt, nd)) return;
    const int ni = nd.cell;
    const long long id = IX(d, nd.k, nd.j, nd.i);
    const long long q1 = IX(d, nd.k1, nd.j1, nd.i1), q2 = IX(d, nd.k2, nd.j2, nd.i2), q3 = IX(d, nd.k3, nd.j3, nd.i3);
    const double sb = nd.d_s, sc = sb + nd.d_i;
    const double sk1 = nd.cr1, sk2 = nd.cr2, sk3 = nd.cr3;
    const double nfx = s.nfx[ni], nfy = s.nfy[ni], nfz = s.nfz[ni];
    double Ua[3] = {0, 0, 0}, Uc[3], Ub[3];
    if (ni >= 0 && s.u) {
        const double *u1 = &s.u[3 * (long long)s.nv1[ni]], *u2 = &s.u[3 * (long long)s.nv2[ni]],
                     *u3 = &s.u[3 * (long long)s.nv3[ni]];
        for (int q = 0; q < 3; q++) Ua[q] = u1[q] * nd.cs1 + u2[q] * nd.cs2 + u3[q] * nd.cs3;
    }
    for (int q = 0; q < 3; q++) Uc[q] = (ucat[3 * q1 + q] * sk1 + ucat[3 * q2 + q] * sk2 + ucat[3 * q3 + q] * sk3);
    if ((!opt.moving && !d.inviscid && nvert[q1] + nvert[q2] + nvert[q3] > 0.1) || nvert[id] > 2.9) {
        Ub[0] = Ub[1] = Ub[2] = 0;
        ustar[id] = 0;
    } else {
        const double u_c = Uc[0] - Ua[0], v_c = Uc[1] - Ua[1], w_c = Uc[2] - Ua[2];
        const double un = u_c * nfx + v_c * nfy + w_c * nfz;
        double ut = u_c - un * nfx, vt = v_c - un * nfy, wt = w_c - un * nfz;
        if (opt.wallfunction == 2) {   // wall_function_roughness_loglaw, wallfunction.c:71-95
            const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
            const double us = rl_utau(ut_mag, sc, 0.01, opt.roughness);
            ustar[id] = us;
            const double um = rl_u(sb, us, opt.roughness);
            if (ut_mag > 1.e-10) ut *= um / ut_mag, vt *= um / ut_mag, wt *= um / ut_mag;
            else ut = vt = wt = 0;
            Ub[0] = ut;
            Ub[1] = vt;
            Ub[2] = wt;
        } else if (opt.wallfunction) {   // wall_function_roughness_a, sediment_transport.c:70-96
            const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
            const double us = hy_utau(1. / d.ren, ut_mag, sc, 0.01, opt.ks);
            ustar[id] = us;
            const double um = hy_u(1. / d.ren, sb, us, opt.ks);
            if (ut_mag > 1.e-10) ut *= um / ut_mag, vt *= um / ut_mag, wt *= um / ut_mag;
            else ut = vt = wt = 0;
            Ub[0] = ut + sb / sc * un * nfx;
            Ub[1] = vt + sb / sc * un * nfy;
            Ub[2] = wt + sb / sc * un * nfz;
        } else if (d.inviscid) {   // freeslip, wallfunction.c:152-166
            Ub[0] = ut + sb / sc * un * nfx;
            Ub[1] = vt + sb / sc * un * nfy;
            Ub[2] = wt + sb / sc * un * nfz;
        } else {            // noslip, wallfunction.c:134-150
            const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
            ustar[id] = sqrt(ut_mag / sc / d.ren);
            Ub[0] = sb / sc * u_c;
            Ub[1] = sb / sc * v_c;
            Ub[2] = sb / sc * w_c;
        }
        Ub[0] += Ua[0], Ub[1] += Ua[1], Ub[2] += Ua[2];
    }
    stage[l] = Ub[0];
    stage[n + l] = Ub[1];
    stage[2 * (long long)n + l] = Ub[2];
    stage[3 * (long long)n + l] = (pf[q1] * sk1 + pf[q2] * sk2 + pf[q3] * sk3);
}

__global__ void __launch_bounds__(256) k_ib_scatter(Dims d, int n_elmt, const vfs_ib_node *__restrict__ nodes, int n,
                                                    const double *__restrict__ stage, double *__restrict__ ucat,
                                                    double *__restrict__ pf)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const vfs_ib_node nd = nodes[l];
    if (!node_ok(d, n_elmt, nd)) return;
    const long long id = IX(d, nd.k, nd.j, nd.i);
    ucat[3 * id] = stage[l];
    ucat[3 * id + 1] = stage[n + l];
    ucat[3 * id + 2] = stage[2 * (long long)n + l];
    pf[id] = stage[3 * (long long)n + l];
}

// ibm.c:3556-3576: per element, the count and the means of ustar^2 and p
// over its nodes, added in list order (one thread an element, the nodes
// staged through LDS)
constexpr int STAT_BLOCK = 256;
__global__ void __launch_bounds__(STAT_BLOCK) k_ib_stats(Dims d, int n_elmt, const vfs_ib_node *__restrict__ nodes,
                                                         int n, const double *__restrict__ ustar,
                                                         const double *__restrict__ stage, int *__restrict__ count,
                                                         double *__restrict__ shear, double *__restrict__ pres)
{
    __shared__ int sc[STAT_BLOCK];
    __shared__ double su[STAT_BLOCK], sp[STAT_BLOCK];
    const int e = blockIdx.x * STAT_BLOCK + threadIdx.x;
    int cnt = 0;
    double sh = 0, pr = 0;
    for (int l0 = 0; l0 < n; l0 += STAT_BLOCK) {
        const int l = l0 + threadIdx.x;
        if (l < n) {
            const vfs_ib_node nd = nodes[l];
            const bool ok = node_ok(d, n_elmt, nd);
            sc[threadIdx.x] = ok ? nd.cell : -1;
            su[threadIdx.x] = ok ? ustar[IX(d, nd.k, nd.j, nd.i)] : 0.0;
            sp[threadIdx.x] = ok ? stage[3 * (long long)n + l] : 0.0;
        }
        __syncthreads();
        const int m = min(STAT_BLOCK, n - l0);
        for (int q = 0; q < m; q++)
            if (sc[q] == e) {
                cnt++;
                sh += su[q] * su[q];
                pr += sp[q];
            }
        __syncthreads();
    }
    if (e >= n_elmt) return;
    if (cnt) {
        sh /= (double)cnt;
        pr /= (double)cnt;
    }
    if (count) count[e] = cnt;
    if (shear) shear[e] = sh;
    if (pres) pres[e] = pr;
}

// ibm.c:3285-3292 Contravariant_Reynolds_stress
__device__ double crs(double uu, double uv, double uw, double vv, double vw, double ww, const double *a, const double *b)
{
    const double A = uu * a[0] * b[0] + vv * a[1] * b[1] + ww * a[2] * b[2] + uv * (a[0] * b[1] + a[1] * b[0]) +
                     uw * (a[0] * b[2] + a[2] * b[0]) + vw * (a[1] * b[2] + a[2] * b[1]);
    const double B = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) * sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    return A / B;
}

// ibm.c:3481-3521: per node, ustar_avg^2 and the three contravariant
// Reynolds stresses from the running sums, into the staging array (four
// blocks of n)
__global__ void __launch_bounds__(256) k_ib_avg_nodes(Dims d, Geom g, IbSurf s, const vfs_ib_node *__restrict__ nodes,
                                                      int n, vfs_ib_options opt, const double *__restrict__ usum,
                                                      const double *__restrict__ u1sum, const double *__restrict__ u2sum,
                                                      double N, double *__restrict__ stage)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const vfs_ib_node nd = nodes[l];
    if (!node_ok(d, s.n_elmt, nd)) return;
    const int ni = nd.cell;
    const long long id = IX(d, nd.k, nd.j, nd.i);
    const long long q1 = IX(d, nd.k1, nd.j1, nd.i1), q2 = IX(d, nd.k2, nd.j2, nd.i2), q3 = IX(d, nd.k3, nd.j3, nd.i3);
    const double sb = nd.d_s, sc = sb + nd.d_i, sk1 = nd.cr1, sk2 = nd.cr2, sk3 = nd.cr3;
    const double nfx = s.nfx[ni], nfy = s.nfy[ni], nfz = s.nfz[ni];
    double Ua[3] = {0, 0, 0}, m[3], m1[3], m2[3];
    if (s.u) {
        const double *u1 = &s.u[3 * (long long)s.nv1[ni]], *u2 = &s.u[3 * (long long)s.nv2[ni]],
                     *u3 = &s.u[3 * (long long)s.nv3[ni]];
        for (int q = 0; q < 3; q++) Ua[q] = u1[q] * nd.cs1 + u2[q] * nd.cs2 + u3[q] * nd.cs3;
    }
    for (int q = 0; q < 3; q++) {
        m[q] = (usum[3 * q1 + q] * sk1 + usum[3 * q2 + q] * sk2 + usum[3 * q3 + q] * sk3) / N;
        m2[q] = (u2sum[3 * q1 + q] * sk1 + u2sum[3 * q2 + q] * sk2 + u2sum[3 * q3 + q] * sk3) / N;
        m1[q] = (u1sum[3 * q1 + q] * sk1 + u1sum[3 * q2 + q] * sk2 + u1sum[3 * q3 + q] * sk3) / N;
    }
    const double u_c = m[0] - Ua[0], v_c = m[1] - Ua[1], w_c = m[2] - Ua[2];
    const double un = u_c * nfx + v_c * nfy + w_c * nfz;
    const double ut = u_c - un * nfx, vt = v_c - un * nfy, wt = w_c - un * nfz;
    const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
    const double us = opt.wallfunction ? hy_utau(1. / d.ren, ut_mag, sc, 0.01, opt.ks) : sqrt(ut_mag / sc / d.ren);
    const double U = m[0], V = m[1], W = m[2];
    const double uu = m2[0] - U * U, vv = m2[1] - V * V, ww = m2[2] - W * W;
    const double uv = m1[0] - U * V, vw = m1[1] - V * W, uw = m1[2] - W * U;
    const double *cs = &g.csi[3 * id], *et = &g.eta[3 * id], *ze = &g.zet[3 * id];
    stage[l] = us * us;
    stage[n + l] = crs(uu, uv, uw, vv, vw, ww, cs, et);
    stage[2 * (long long)n + l] = crs(uu, uv, uw, vv, vw, ww, et, ze);
    stage[3 * (long long)n + l] = crs(uu, uv, uw, vv, vw, ww, cs, ze);
}

// ibm.c:3556-3576 for the averaging statistics: per element, the means of
// the four staged values over its nodes, added in list order
__global__ void __launch_bounds__(STAT_BLOCK) k_ib_avg_stats(Dims d, int n_elmt, const vfs_ib_node *__restrict__ nodes,
                                                             int n, const double *__restrict__ stage,
                                                             double *__restrict__ o0, double *__restrict__ o1,
                                                             double *__restrict__ o2, double *__restrict__ o3)
{
    __shared__ int sc[STAT_BLOCK];
    __shared__ double sv[4][STAT_BLOCK];
    const int e = blockIdx.x * STAT_BLOCK + threadIdx.x;
    int cnt = 0;
    double a[4] = {0, 0, 0, 0};
    for (int l0 = 0; l0 < n; l0 += STAT_BLOCK) {
        const int l = l0 + threadIdx.x;
        if (l < n) {
            const vfs_ib_node nd = nodes[l];
            const bool ok = node_ok(d, n_elmt, nd);
            sc[threadIdx.x] = ok ? nd.cell : -1;
            for (int q = 0; q < 4; q++) sv[q][threadIdx.x] = ok ? stage[(long long)q * n + l] : 0.0;
        }
        __syncthreads();
        const int m = min(STAT_BLOCK, n - l0);
        for (int q = 0; q < m; q++)
            if (sc[q] == e) {
                cnt++;
                for (int w = 0; w < 4; w++) a[w] += sv[w][q];
            }
        __syncthreads();
    }
    if (e >= n_elmt) return;
    if (cnt)
        for (int w = 0; w < 4; w++) a[w] /= (double)cnt;
    o0[e] = a[0];
    o1[e] = a[1];
    o2[e] = a[2];
    o3[e] = a[3];
}

// ibm.c:3619-3679: the fluxes of the faces of every nvert == 1 cell from the
// averaged new velocity (immersed != 3: factor 1)
__global__ void __launch_bounds__(256) k_ib_faces(Dims d, Geom g, const double *__restrict__ nvert,
                                                  const double *__restrict__ ucat, double *__restrict__ ucont)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || i > d.mx - 2 || j < 1 || j > d.my - 2 || k < 1 || k > d.mz - 2) return;
    const long long ii = t + 1, jj = t + d.mx, kk = t + (long long)d.mx * d.my;
    const int *bt = d.bc;
    const double u0 = ucat[3 * t], v0 = ucat[3 * t + 1], w0 = ucat[3 * t + 2];
    const bool self = (int)(nvert[t] + 0.5) == 1, ni = (int)(nvert[ii] + 0.5) == 1,
               nj = (int)(nvert[jj] + 0.5) == 1, nk = (int)(nvert[kk] + 0.5) == 1;
    double ux = ucont[3 * t], uy = ucont[3 * t + 1], uz = ucont[3 * t + 2];
    // each flux is the same average whichever cell asks for it; the wall
    // rule of the asking cell applies
    double fx = 0, fy = 0, fz = 0;
    if (self || ni) {
        const double ucx = (u0 + ucat[3 * ii]) * 0.5, ucy = (v0 + ucat[3 * ii + 1]) * 0.5,
                     ucz = (w0 + ucat[3 * ii + 2]) * 0.5;
        fx = (ucx * g.icsi[3 * t] + ucy * g.icsi[3 * t + 1] + ucz * g.icsi[3 * t + 2]) * 1.0;
    }
    if (self || nj) {
        const double ucx = (ucat[3 * jj] + u0) * 0.5, ucy = (ucat[3 * jj + 1] + v0) * 0.5,
                     ucz = (ucat[3 * jj + 2] + w0) * 0.5;
        fy = (ucx * g.jeta[3 * t] + ucy * g.jeta[3 * t + 1] + ucz * g.jeta[3 * t + 2]) * 1.0;
    }
    if (self || nk) {
        const double ucx = (ucat[3 * kk] + u0) * 0.5, ucy = (ucat[3 * kk + 1] + v0) * 0.5,
                     ucz = (ucat[3 * kk + 2] + w0) * 0.5;
        fz = (ucx * g.kzet[3 * t] + ucy * g.kzet[3 * t + 1] + ucz * g.kzet[3 * t + 2]) * 1.0;
    }
    if (self) {
        ux = ((bt[0] == -1 || bt[0] == -2) && i == 1) ? 0 : fx;
        uy = ((bt[2] == -1 || bt[2] == -2) && j == 1) ? 0 : fy;
        uz = ((bt[4] == -1 || bt[4] == -2) && k == 1) ? 0 : fz;
    }
    if (ni) ux = ((bt[1] == -1 || bt[1] == -2) && i == d.mx - 3) ? 0 : fx;
    if (nj) uy = ((bt[3] == -1 || bt[3] == -2) && j == d.my - 3) ? 0 : fy;
    if (nk) uz = ((bt[5] == -1 || bt[5] == -2) && k == d.mz - 3) ? 0 : fz;
    ucont[3 * t] = ux;
    ucont[3 * t + 1] = uy;
    ucont[3 * t + 2] = uz;
}

// ---- the marking of ibm_search_advanced (ibm.c:110-360) ----------------
constexpr int NBIN = 40;        // ncx = ncy = ncz (ibm.c:104)
constexpr int NTRY = 32;        // rays cast before a cell is reported undecided
constexpr int MAXHIT = 32;      // crossings kept per ray for the coincidence test

struct BodyBox {
    double xmin, ymin, zmin, xmax, ymax, zmax, dcx, dcy, dcz;
};

// compgeom.c:17-51 intsect_triangle (Moller-Trumbore, EPSILON 1e-8)
__device__ int intsect(const double o[3], const double dir[3], const double v0[3], const double v1[3],
                       const double v2[3], double &t)
{
    double e1[3], e2[3], tv[3], pv[3], qv[3];
    for (int q = 0; q < 3; q++) e1[q] = v1[q] - v0[q], e2[q] = v2[q] - v0[q];
    pv[0] = dir[1] * e2[2] - dir[2] * e2[1];
    pv[1] = dir[2] * e2[0] - dir[0] * e2[2];
    pv[2] = dir[0] * e2[1] - dir[1] * e2[0];
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det > -1e-8 && det < 1e-8) return 0;
    const double inv = 1.0 / det;
    for (int q = 0; q < 3; q++) tv[q] = o[q] - v0[q];
    const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (u < 0.0 || u > 1.0) return 0;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
    qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
    qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    const double v = (dir[0] * qv[0] + dir[1] * qv[1] + dir[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return 0;
    t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    return 1;
}

// point_cell_advanced (ibm.c:1068-1183) for every interior cell whose centre
// lies strictly inside the widened box: the elements listed in the bins
// (ic, jc, k >= kc) are those whose own bin range holds that column, tested
// once each.  The result (the parity of the crossings, or a coincidence that
// asks for the next ray) does not depend on the order the elements are
// tested in.  The elements whose bin range holds column (ic, jc) are listed
// per column (cstart / celem, element order), so a cell tests those only.
// reset clears nvert < 5 first (the first body of a search).
__global__ void __launch_bounds__(256) k_body_inside(Dims d, const double *__restrict__ cent, IbSurf s,
                                                     const int *__restrict__ bins, const int *__restrict__ cstart,
                                                     const int *__restrict__ celem, BodyBox b,
                                                     const double *__restrict__ jitter, int reset,
                                                     double *__restrict__ nvert, int *__restrict__ fail)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || i > d.mx - 2 || j < 1 || j > d.my - 2 || k < 1 || k > d.mz - 2) return;
    double nv = nvert[t];
    if (reset && nv < 5.) nv = 0;
    const double p[3] = {cent[3 * t], cent[3 * t + 1], cent[3 * t + 2]};
    if (p[0] > b.xmin && p[0] < b.xmax && p[1] > b.ymin && p[1] < b.ymax && p[2] > b.zmin && p[2] < b.zmax) {
        const int ic = (int)floor((p[0] - b.xmin) / b.dcx), jc = (int)floor((p[1] - b.ymin) / b.dcy),
                  kc = (int)floor((p[2] - b.zmin) / b.dcz);
        int val = -1;
        for (int seed = 1; seed <= NTRY && val < 0; seed++) {
            const double xpc = b.dcx * (ic + 0.5) + b.xmin, ypc = b.dcy * (jc + 0.5) + b.ymin;
            const double sj = jitter[seed - 1];
            double dir[3] = {xpc + sj * b.dcx - p[0], ypc + sj * b.dcy - p[1], b.zmax + 0.2 - p[2]};
            const double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
            dir[0] /= len, dir[1] /= len, dir[2] /= len;
            double dt[MAXHIT];
            int de[MAXHIT];
            int nintp = 0, sing = 0;
            const int col = ic * NBIN + jc;
            for (int qe = cstart[col]; qe < cstart[col + 1] && !sing; qe++) {
                const int e = celem[qe];
                const int *r = &bins[6 * e];
                if (!(r[0] <= ic && ic < r[1] && r[2] <= jc && jc < r[3] && (r[4] > kc ? r[4] : kc) < r[5])) continue;
                double v0[3], v1[3], v2[3], tt;
                vertex(s, s.nv1[e], v0);
                vertex(s, s.nv2[e], v1);
                vertex(s, s.nv3[e], v2);
                if (!(intsect(p, dir, v0, v1, v2, tt) > 0 && tt > 0)) continue;
                if (nintp == MAXHIT) {
                    sing = 2;
                    break;
                }
                for (int m = 0; m < nintp; m++) {
                    const int f = de[m];
                    const double ndotn = s.nfx[f] * s.nfx[e] + s.nfy[f] * s.nfy[e] + s.nfz[f] * s.nfz[e];
                    if (fabs(tt - dt[m]) < 1e-8 && ndotn > -0.97) sing = 1;
                }
                dt[nintp] = tt;
                de[nintp] = e;
                nintp++;
            }
            if (sing == 2) break;
            if (!sing) val = nintp % 2 ? 4 : 0;
        }
        if (val < 0) atomicAdd(fail, 1);
        else nv = nv > val ? nv : (double)val;
        if (nv < 0) nv = 0;
    }
    nvert[t] = nv;
}

// compgeom.c:301-370 ISLineTriangleIntp: does segment p1p2 cut element e
__device__ bool line_cuts(const IbSurf &s, const double p1[3], const double p2[3], int e)
{
    double f1[3], f2[3], f3[3];
    vertex(s, s.nv1[e], f1);
    vertex(s, s.nv2[e], f2);
    vertex(s, s.nv3[e], f3);
    const double nfx = s.nfx[e], nfy = s.nfy[e], nfz = s.nfz[e];
    const double tf1 = (p1[0] - f1[0]) * nfx + (p1[1] - f1[1]) * nfy + (p1[2] - f1[2]) * nfz;
    const double tf2 = (p2[0] - f1[0]) * nfx + (p2[1] - f1[1]) * nfy + (p2[2] - f1[2]) * nfz;
    if (!(tf1 * tf2 < 0)) return false;
    const double dx1 = p1[0] - f1[0], dy1 = p1[1] - f1[1], dz1 = p1[2] - f1[2];
    const double nftx = p2[0] - p1[0], nfty = p2[1] - p1[1], nftz = p2[2] - p1[2];
    const double dx2 = f2[0] - f1[0], dy2 = f2[1] - f1[1], dz2 = f2[2] - f1[2];
    const double dx3 = f3[0] - f1[0], dy3 = f3[1] - f1[1], dz3 = f3[2] - f1[2];
    const double d = -(dx1 * (dy2 * dz3 - dz2 * dy3) - dy1 * (dx2 * dz3 - dz2 * dx3) + dz1 * (dx2 * dy3 - dy2 * dx3)) /
                     (nftx * (dy2 * dz3 - dz2 * dy3) - nfty * (dx2 * dz3 - dz2 * dx3) + nftz * (dx2 * dy3 - dy2 * dx3));
    const double pint[3] = {p1[0] + d * nftx, p1[1] + d * nfty, p1[2] + d * nftz};
    if (!(d > 0 && d < 1.)) return false;
    return point_in_triangle(pint, f1, f2, f3, nfx, nfy, nfz) == 1;
}

// -thin (ibm.c:281-317), first pass: whether a segment from the cell centre
// to its +i, +j, +k and +ijk neighbours cuts an element of the bins around
// it (point_cell_thin, ibm.c:1186-1244), for cells in the box whose 2x2x2
// box still has a cell below 0.5.  The marks go in a second pass: a cell that
// the sequential loop would skip has only cells already marked in its box,
// so marking from the unchanged mask gives the same result.
__global__ void __launch_bounds__(256) k_thin_cut(Dims d, const double *__restrict__ cent, IbSurf s,
                                                  const int *__restrict__ bins, BodyBox b,
                                                  const double *__restrict__ nvert, int *__restrict__ cut)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    int c = 0;
    const double p[3] = {cent[3 * t], cent[3 * t + 1], cent[3 * t + 2]};
    if (i >= 1 && i <= d.mx - 2 && j >= 1 && j <= d.my - 2 && k >= 1 && k <= d.mz - 2 && p[0] > b.xmin &&
        p[0] < b.xmax && p[1] > b.ymin && p[1] < b.ymax && p[2] > b.zmin && p[2] < b.zmax) {
        bool open = false;
        for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++) open |= nvert[IX(d, k + dk, j + dj, i + di)] < 0.5;
        if (open) {
            const int ic = (int)floor((p[0] - b.xmin) / b.dcx), jc = (int)floor((p[1] - b.ymin) / b.dcy),
                      kc = (int)floor((p[2] - b.zmin) / b.dcz);
            const int is = ic - 1 > 0 ? ic - 1 : 0, js = jc - 1 > 0 ? jc - 1 : 0, ks = kc - 1 > 0 ? kc - 1 : 0;
            const int ie = ic + 2 < NBIN ? ic + 2 : NBIN, je = jc + 2 < NBIN ? jc + 2 : NBIN,
                      ke = kc + 2 < NBIN ? kc + 2 : NBIN;
            const long long nb[4] = {IX(d, k, j, i + 1), IX(d, k, j + 1, i), IX(d, k + 1, j, i), IX(d, k + 1, j + 1, i + 1)};
            for (int e = 0; e < s.n_elmt && !c; e++) {
                const int *r = &bins[6 * e];
                if (r[0] >= ie || r[1] <= is || r[2] >= je || r[3] <= js || r[4] >= ke || r[5] <= ks) continue;
                for (int w = 0; w < 4 && !c; w++) {
                    const double q[3] = {cent[3 * nb[w]], cent[3 * nb[w] + 1], cent[3 * nb[w] + 2]};
                    if (line_cuts(s, p, q, e)) c = 1;
                }
            }
        }
    }
    cut[t] = c;
}

// second pass: 2 where the cell is below 0.5 and a cut cell's box holds it
__global__ void __launch_bounds__(256) k_thin_mark(Dims d, const int *__restrict__ cut, double *__restrict__ nvert)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    if (!(nvert[t] < 0.5)) return;
    const int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    bool m = false;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                const int kk = k - dk, jj = j - dj, ii = i - di;
                if (kk >= 1 && jj >= 1 && ii >= 1) m |= cut[IX(d, kk, jj, ii)] != 0;
            }
    if (m) nvert[t] = 2.;
}

int check_surface(const vfs_ib_surface *s)
{
    VFS_REQUIRE(s && s->n_v > 0 && s->n_elmt > 0, "the body needs vertices and elements");
    VFS_REQUIRE(s->x_bp && s->y_bp && s->z_bp && s->nv1 && s->nv2 && s->nv3 && s->nf_x && s->nf_y && s->nf_z,
                "bad surface arrays");
    return 0;
}

}   // namespace

extern "C" int vfs_ib_search(vfs_ctx *c, const vfs_ib_surface *surf, vfs_ib_node *nodes, int cap, int *n_nodes)
{
    const Dims d = c->d;
    if (int rc = check_surface(surf)) return rc;
    VFS_REQUIRE(n_nodes && cap >= 0 && (cap == 0 || nodes), "bad node list");
    const IbSurf s = to_surf(surf);
    const int nb = (int)grid_for(d.P, LIST_BLOCK);
    // the element bounding spheres, the per-block node counts and the fail flag
    const size_t bytes = (size_t)s.n_elmt * 4 * sizeof(double) + (size_t)nb * 2 * sizeof(int) + sizeof(int);
    char *buf = nullptr;
    VFS_CHECK_HIP(hipMalloc(&buf, bytes));
    double *qv = (double *)buf, *rad = qv + 3 * (size_t)s.n_elmt;
    int *cnt = (int *)(rad + s.n_elmt), *off = cnt + nb, *fail = off + nb;
    std::vector<int> h(nb);
    int total = 0, nfail = 0;
    hipError_t e = hipMemsetAsync(fail, 0, sizeof(int), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ib_bound, dim3(grid_for(s.n_elmt, 256)), dim3(256), 0, c->stream, s, qv, rad);
        hipLaunchKernelGGL(k_ib_count, dim3(nb), dim3(LIST_BLOCK), 0, c->stream, d, c->nvert, cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), cnt, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) {
        for (int b = 0; b < nb; b++) {   // exclusive prefix of the block counts
            const int v = h[b];
            h[b] = total;
            total += v;
        }
        if (total <= cap && total > 0) {
            e = hipMemcpyAsync(off, h.data(), nb * sizeof(int), hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_ib_place, dim3(nb), dim3(LIST_BLOCK), 0, c->stream, d, c->nvert, off, total, nodes);
                hipLaunchKernelGGL(k_ib_search, dim3(total), dim3(64), 0, c->stream, d, c->g, s, qv, rad,
                                   total, nodes, fail);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipMemcpyAsync(&nfail, fail, sizeof(int), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        }
    }
    hipFree(buf);
    VFS_CHECK_HIP(e);
    *n_nodes = total;
    if (total > cap) {
        vfs_set_error("vfs_ib_search: " + std::to_string(total) + " IB nodes do not fit the list of " +
                      std::to_string(cap));
        return -4;
    }
    if (nfail) {
        vfs_set_error("vfs_ib_search: " + std::to_string(nfail) +
                      " IB nodes have no nearest element or no interception point");
        return -3;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// vfs_ib_search_noalloc: the same node list with its scratch (the element
// bounding spheres, the per-block node counts turned into offsets in place,
// the node total and the failure count) in a static device array instead of
// a per-call allocation.  Each device holds its own copy of the array until
// the process ends; calls that overlap on one device share it and race.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t IB_SCRATCH_BYTES = size_t(8) << 20;
constexpr int IB_SCRATCH_DEVICES = 64;

static __device__ __align__(256) char scratch[IB_SCRATCH_BYTES];

// the array's device address on each device (the host-side name of the array
// is not a device address), written once under the mutex
char *g_scratch_dev[IB_SCRATCH_DEVICES];
std::mutex g_scratch_mu;

// exclusive scan of cnt[0..nb) in place, with the total at *total (one
// workgroup of 256 in chunks of 256)
__global__ void __launch_bounds__(256) k_ib_scan(int nb, int *cnt, int *total)
{
    __shared__ int sh[256];
    int carry = 0;
    for (int base = 0; base < nb; base += 256) {
        const int v = base + (int)threadIdx.x < nb ? cnt[base + threadIdx.x] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int add = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (base + (int)threadIdx.x < nb) cnt[base + threadIdx.x] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        carry += sh[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_ib_zero_ints(int n, int *p)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

}   // namespace

// Called by vfs_ctx_create with the context's device current: records that
// device's address of the static scratch once.
int vfs_ib_scratch_setup(int device)
{
    VFS_REQUIRE(device >= 0 && device < IB_SCRATCH_DEVICES, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return 0;
    void *p = nullptr;
    VFS_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = (char *)p;
    return 0;
}

extern "C" int vfs_ib_search_noalloc(vfs_ctx *c, const vfs_ib_surface *surf, vfs_ib_node *nodes, int cap,
                                     int *n_nodes)
{
    const Dims d = c->d;
    if (int rc = check_surface(surf)) return rc;
    VFS_REQUIRE(n_nodes && cap >= 0 && (cap == 0 || nodes), "bad node list");
    char *buf = nullptr;
    if (c->device >= 0 && c->device < IB_SCRATCH_DEVICES) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        buf = g_scratch_dev[c->device];
    }
    VFS_REQUIRE(buf, "vfs_ib_search_noalloc: the context's device has no scratch address");
    const IbSurf s = to_surf(surf);
    const int nb = (int)grid_for(d.P, LIST_BLOCK);
    // layout: qv (3 n_elmt doubles), rad (n_elmt doubles), cnt (nb ints), total, fail
    const size_t bytes = (size_t)s.n_elmt * 4 * sizeof(double) + ((size_t)nb + 2) * sizeof(int);
    if (bytes > IB_SCRATCH_BYTES) {
        vfs_set_error("vfs_ib_search_noalloc: " + std::to_string(s.n_elmt) + " elements and " + std::to_string(d.P) +
                      " grid points need " + std::to_string(bytes) + " bytes of scratch; the static scratch holds " +
                      std::to_string(IB_SCRATCH_BYTES) + " (32 bytes an element plus 4 bytes per 256 points)");
        return -5;
    }
    double *qv = (double *)buf, *rad = qv + 3 * (size_t)s.n_elmt;
    int *cnt = (int *)(rad + s.n_elmt), *total_d = cnt + nb, *fail = total_d + 1;
    hipLaunchKernelGGL(k_ib_zero_ints, dim3(1), dim3(64), 0, c->stream, 2, total_d);
    hipLaunchKernelGGL(k_ib_bound, dim3(grid_for(s.n_elmt, 256)), dim3(256), 0, c->stream, s, qv, rad);
    hipLaunchKernelGGL(k_ib_count, dim3(nb), dim3(LIST_BLOCK), 0, c->stream, d, c->nvert, cnt);
    hipLaunchKernelGGL(k_ib_scan, dim3(1), dim3(256), 0, c->stream, nb, cnt, total_d);
    VFS_CHECK_LAUNCH();
    int total = 0, nfail = 0;
    VFS_CHECK_HIP(hipMemcpyAsync(&total, total_d, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    *n_nodes = total;
    if (total > cap) {
        vfs_set_error("vfs_ib_search_noalloc: " + std::to_string(total) + " IB nodes do not fit the list of " +
                      std::to_string(cap));
        return -4;
    }
    if (total > 0) {
        hipLaunchKernelGGL(k_ib_place, dim3(nb), dim3(LIST_BLOCK), 0, c->stream, d, c->nvert, cnt, total, nodes);
        hipLaunchKernelGGL(k_ib_search, dim3(total), dim3(64), 0, c->stream, d, c->g, s, qv, rad, total,
                           nodes, fail);
        VFS_CHECK_LAUNCH();
        VFS_CHECK_HIP(hipMemcpyAsync(&nfail, fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    }
    if (nfail) {
        vfs_set_error("vfs_ib_search_noalloc: " + std::to_string(nfail) +
                      " IB nodes have no nearest element or no interception point");
        return -3;
    }
    return 0;
}

extern "C" int vfs_ib_interpolate(vfs_ctx *c, const vfs_ib_surface *surf, const vfs_ib_node *nodes, int n,
                                  double *ucont, double *ucat, double *p, double *ustar, int *count, double *shear,
                                  double *pressure)
{
    const vfs_ib_options o = {0, 0, 0.0};
    return vfs_ib_interpolate_opt(c, surf, nodes, n, &o, ucont, ucat, p, ustar, count, shear, pressure);
}

extern "C" int vfs_ib_interpolate_opt(vfs_ctx *c, const vfs_ib_surface *surf, const vfs_ib_node *nodes, int n,
                                      const vfs_ib_options *opt, double *ucont, double *ucat, double *p, double *ustar,
                                      int *count, double *shear, double *pressure)
{
    const Dims d = c->d;
    VFS_REQUIRE(opt && (opt->moving == 0 || opt->moving == 1) && opt->wallfunction >= 0 && opt->wallfunction <= 2,
                "bad IB options");
    VFS_REQUIRE(opt->wallfunction != 1 || opt->ks > 0, "the IB wall model needs a roughness ks > 0");
    VFS_REQUIRE(opt->wallfunction != 2 || opt->roughness > 0, "the rough log law needs a roughness > 0");
    if (int rc = check_surface(surf)) return rc;
    VFS_REQUIRE(n >= 0 && (n == 0 || nodes) && ucont && ucat && p && ustar, "bad arguments");
    VFS_REQUIRE(4 * (long long)n <= 3 * d.P, "more IB nodes than cells");
    const IbSurf s = to_surf(surf);
    double *stage = c->tmp3;   // 4n doubles
    if (n > 0) {
        hipLaunchKernelGGL(k_ib_nodes, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d, c->nvert, s, nodes, n, ucat, p,
                           *opt, ustar, stage);
        hipLaunchKernelGGL(k_ib_scatter, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d, s.n_elmt, nodes, n, stage,
                           ucat, p);
        VFS_CHECK_LAUNCH();
    }
    if (count || shear || pressure) {
        hipLaunchKernelGGL(k_ib_stats, dim3(grid_for(s.n_elmt, STAT_BLOCK)), dim3(STAT_BLOCK), 0, c->stream, d,
                           s.n_elmt, nodes, n, ustar, stage, count, shear, pressure);
        VFS_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ib_faces, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, ucat, ucont);
    VFS_CHECK_LAUNCH();
    return vfs_contra2cart(c, ucont, ucat);
}

// the first rand() after srand(seed) of glibc, scaled to [-0.5, 0.5) as
// randomdirection does (ibm.c:52-84), from a private generator state so the
// caller's rand() sequence is left alone (srand/rand are random()'s TYPE_3
// generator with its 128-byte state)
extern "C" int vfs_ib_ray_jitter(int seed, double *s)
{
    VFS_REQUIRE(s, "bad output");
    struct random_data rd = {};
    char state[128];
    int32_t v = 0;
    VFS_REQUIRE(initstate_r((unsigned)seed, state, sizeof state, &rd) == 0 && random_r(&rd, &v) == 0,
                "random_r failed");
    *s = v / ((double)RAND_MAX + 1) - 0.5;
    return 0;
}

extern "C" int vfs_nvert_body(vfs_ctx *c, const vfs_ib_surface *surf, int reset, double *nvert_out)
{
    return vfs_nvert_body_thin(c, surf, reset, 0, nvert_out);
}

extern "C" int vfs_nvert_body_thin(vfs_ctx *c, const vfs_ib_surface *surf, int reset, int thin, double *nvert_out)
{
    c->nvert_version++;
    const Dims d = c->d;
    if (int rc = check_surface(surf)) return rc;
    const IbSurf s = to_surf(surf);
    // the widened box and every element's bin range, on the host (ibm.c:135-209)
    std::vector<double> x(s.n_v), y(s.n_v), z(s.n_v);
    std::vector<int> n1(s.n_elmt), n2(s.n_elmt), n3(s.n_elmt);
    VFS_CHECK_HIP(hipMemcpy(x.data(), s.x, s.n_v * sizeof(double), hipMemcpyDeviceToHost));
    VFS_CHECK_HIP(hipMemcpy(y.data(), s.y, s.n_v * sizeof(double), hipMemcpyDeviceToHost));
    VFS_CHECK_HIP(hipMemcpy(z.data(), s.z, s.n_v * sizeof(double), hipMemcpyDeviceToHost));
    VFS_CHECK_HIP(hipMemcpy(n1.data(), s.nv1, s.n_elmt * sizeof(int), hipMemcpyDeviceToHost));
    VFS_CHECK_HIP(hipMemcpy(n2.data(), s.nv2, s.n_elmt * sizeof(int), hipMemcpyDeviceToHost));
    VFS_CHECK_HIP(hipMemcpy(n3.data(), s.nv3, s.n_elmt * sizeof(int), hipMemcpyDeviceToHost));
    BodyBox b;
    b.xmin = b.ymin = b.zmin = 1.e23;
    b.xmax = b.ymax = b.zmax = -1.e23;
    for (int v = 0; v < s.n_v; v++) {
        VFS_REQUIRE(std::isfinite(x[v]) && std::isfinite(y[v]) && std::isfinite(z[v]), "vertex coordinates must be finite");
        b.xmin = std::fmin(b.xmin, x[v]); b.xmax = std::fmax(b.xmax, x[v]);
        b.ymin = std::fmin(b.ymin, y[v]); b.ymax = std::fmax(b.ymax, y[v]);
        b.zmin = std::fmin(b.zmin, z[v]); b.zmax = std::fmax(b.zmax, z[v]);
    }
    b.xmin -= 0.05; b.xmax += 0.05;
    b.ymin -= 0.05; b.ymax += 0.05;
    b.zmin -= 0.05; b.zmax += 0.05;
    b.dcx = (b.xmax - b.xmin) / (NBIN - 1.);
    b.dcy = (b.ymax - b.ymin) / (NBIN - 1.);
    b.dcz = (b.zmax - b.zmin) / (NBIN - 1.);
    std::vector<int> bins(6 * (size_t)s.n_elmt);
    for (int e = 0; e < s.n_elmt; e++) {
        const int a = n1[e], bb = n2[e], cc = n3[e];
        VFS_REQUIRE(a >= 0 && a < s.n_v && bb >= 0 && bb < s.n_v && cc >= 0 && cc < s.n_v, "element vertex out of range");
        const double x0 = std::fmin(std::fmin(x[a], x[bb]), x[cc]), x1 = std::fmax(std::fmax(x[a], x[bb]), x[cc]);
        const double y0 = std::fmin(std::fmin(y[a], y[bb]), y[cc]), y1 = std::fmax(std::fmax(y[a], y[bb]), y[cc]);
        const double z0 = std::fmin(std::fmin(z[a], z[bb]), z[cc]), z1 = std::fmax(std::fmax(z[a], z[bb]), z[cc]);
        int r[6] = {(int)std::floor((x0 - b.xmin) / b.dcx), (int)std::floor((x1 - b.xmin) / b.dcx) + 1,
                    (int)std::floor((y0 - b.ymin) / b.dcy), (int)std::floor((y1 - b.ymin) / b.dcy) + 1,
                    (int)std::floor((z0 - b.zmin) / b.dcz), (int)std::floor((z1 - b.zmin) / b.dcz) + 1};
        for (int q = 0; q < 6; q += 2) {
            r[q] = r[q] < 0 ? 0 : r[q];
            r[q + 1] = r[q + 1] > NBIN ? NBIN : r[q + 1];
        }
        for (int q = 0; q < 6; q++) bins[6 * (size_t)e + q] = r[q];
    }
    // the elements of each (ic, jc) column, in element order
    std::vector<int> cstart(NBIN * NBIN + 1, 0), celem;
    for (int e = 0; e < s.n_elmt; e++)
        for (int ic = bins[6 * (size_t)e]; ic < bins[6 * (size_t)e + 1]; ic++)
            for (int jc = bins[6 * (size_t)e + 2]; jc < bins[6 * (size_t)e + 3]; jc++) cstart[ic * NBIN + jc + 1]++;
    for (int q = 0; q < NBIN * NBIN; q++) cstart[q + 1] += cstart[q];
    celem.resize(std::max(cstart[NBIN * NBIN], 1));
    {
        std::vector<int> at(cstart.begin(), cstart.end() - 1);
        for (int e = 0; e < s.n_elmt; e++)
            for (int ic = bins[6 * (size_t)e]; ic < bins[6 * (size_t)e + 1]; ic++)
                for (int jc = bins[6 * (size_t)e + 2]; jc < bins[6 * (size_t)e + 3]; jc++) celem[at[ic * NBIN + jc]++] = e;
    }
    double jit[NTRY];
    for (int q = 0; q < NTRY; q++)
        if (vfs_ib_ray_jitter(q + 1, &jit[q])) return -1;
    const size_t bytes = bins.size() * sizeof(int) + sizeof jit + sizeof(int) + (cstart.size() + celem.size()) * sizeof(int);
    char *buf = nullptr;
    VFS_CHECK_HIP(hipMalloc(&buf, bytes));
    int *dbins = (int *)buf;
    double *djit = (double *)(buf + bins.size() * sizeof(int));   // 24 bytes an element: 8-aligned
    int *fail = (int *)(djit + NTRY);
    int *dcs = fail + 1, *dce = dcs + cstart.size();
    int nfail = 0;
    hipError_t e = hipMemcpyAsync(dbins, bins.data(), bins.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(djit, jit, sizeof jit, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dcs, cstart.data(), cstart.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dce, celem.data(), celem.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(fail, 0, sizeof(int), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_body_inside, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g.cent, s, dbins,
                           dcs, dce, b, djit, reset ? 1 : 0, c->nvert, fail);
        e = hipGetLastError();
    }
    if (e == hipSuccess && thin) {
        int *cut = (int *)c->tmp3;   // P ints (tmp3 holds 3P doubles)
        hipLaunchKernelGGL(k_thin_cut, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g.cent, s, dbins, b,
                           c->nvert, cut);
        hipLaunchKernelGGL(k_thin_mark, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, cut, c->nvert);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&nfail, fail, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(buf);
    VFS_CHECK_HIP(e);
    if (nfail) {
        vfs_set_error("vfs_nvert_body: " + std::to_string(nfail) + " cells found no clear ray in " +
                      std::to_string(NTRY) + " tries");
        return -3;
    }
    if (vfs_mask_dilate_relabel(c)) return -1;
    c->nvert_f32 = 1;   // small integers
    if (nvert_out)
        VFS_CHECK_HIP(hipMemcpyAsync(nvert_out, c->nvert, (size_t)d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

extern "C" int vfs_ib_averages(vfs_ctx *c, const vfs_ib_surface *surf, const vfs_ib_node *nodes, int n,
                               const vfs_ib_options *opt, const double *usum, const double *u1sum,
                               const double *u2sum, double N, double *mean_shear, double *rs1, double *rs2,
                               double *rs3)
{
    const Dims d = c->d;
    if (int rc = check_surface(surf)) return rc;
    VFS_REQUIRE(opt && n >= 0 && (n == 0 || nodes) && usum && u1sum && u2sum && mean_shear && rs1 && rs2 && rs3,
                "bad arguments");
    VFS_REQUIRE(!opt->wallfunction || opt->ks > 0, "the IB wall model needs a roughness ks > 0");
    VFS_REQUIRE(opt->wallfunction != 2 || opt->roughness > 0, "the rough log law needs a roughness > 0");
    VFS_REQUIRE(4 * (long long)n <= 3 * d.P, "more IB nodes than cells");
    const IbSurf s = to_surf(surf);
    double *stage = c->tmp3;
    if (n > 0) {
        hipLaunchKernelGGL(k_ib_avg_nodes, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d, c->g, s, nodes, n, *opt,
                           usum, u1sum, u2sum, N, stage);
        VFS_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ib_avg_stats, dim3(grid_for(s.n_elmt, STAT_BLOCK)), dim3(STAT_BLOCK), 0, c->stream, d,
                       s.n_elmt, nodes, n, stage, mean_shear, rs1, rs2, rs3);
    VFS_CHECK_LAUNCH();
    return 0;
}
