// Formfunction_2 (momentum.c:456-1979) with the face fluxes computed from
// LDS-staged planes.
//
// The face fluxes dominate the right-hand side: each face reads about 54 ucat
// values and 10 nvert values on a 2x4x3 neighbourhood.  Read from global
// memory as AoS doubles, those loads are bound by the L1/TA rate long before
// HBM.  Here a workgroup owns a 64(i) x 8(j) column of cells and marches in
// k.  It keeps four k-planes of ucat (SoA, FP64) and nvert (FP32, exact for
// the small-integer masks) with a one-cell i/j halo (two on the high i/j
// side) in LDS: planes k-1..k+2 are what the i-, j- and k-faces of plane k
// read.  The next plane is prefetched into registers while the current one
// is computed.  On the tile planes whose box holds no marked cell (a flag per
// tile and plane, recomputed when the mask changes) the faces take the
// mask-free specialisation of face_flux_t.
//
// Every expression is the one in vfs_stencil.h, so results are bit-identical
// to the reference-structured pass k_ff2_faces.  The faces on the low
// boundary planes (i = 0, j = 0, k = 0) and, in periodic directions, the last
// faces, whose 4th-order stencil wraps around the seam, are computed by a
// small separate kernel.  The cell divergence marches in k with the k-face
// fluxes in registers, and the projection runs as in the reference-structured
// path.
#include "vfs_internal.h"
#include "vfs_stencil.h"

int vfs_ff2_passes(vfs_ctx *c, const double *ucont, const double *ucat, const double *nu_t, double *rhs, double scale);
int vfs_ff2_rhs_only(vfs_ctx *c, double *rhs, double scale);

// Shape parameters (tuned on MI355X).
#ifndef VFS_FACES_KCHUNK
#define VFS_FACES_KCHUNK 32      // k-planes per workgroup of the face kernel
#endif
#ifndef VFS_FACES_WAVES
#define VFS_FACES_WAVES 2        // target waves per SIMD of the face kernel
#endif
#ifndef VFS_CELLS_KCHUNK
#define VFS_CELLS_KCHUNK 16
#endif

namespace {

constexpr int TI = 64;          // cells per workgroup row (one wavefront)
#ifndef VFS_FACES_TJ
#define VFS_FACES_TJ 8
#endif
constexpr int TJ = VFS_FACES_TJ; // rows per workgroup
constexpr int BI = TI + 3;      // staged columns: i0-1 .. i0+TI+1
constexpr int BJ = TJ + 3;      // staged rows:    j0-1 .. j0+TJ+1
constexpr int BP = BI * BJ;     // points per staged plane
constexpr int NT = TI * TJ;     // threads per face group
constexpr int NTB = NT;         // threads per workgroup
constexpr int NU = (3 * BP + NTB - 1) / NTB;   // ucat doubles per thread per plane
constexpr int NN = (BP + NTB - 1) / NTB;       // nvert values per thread per plane

// staged planes (one set per workgroup): ucat [4][3][BJ][BI], nvert [4][BJ][BI]
__shared__ double sU[4 * 3 * BP];
__shared__ float sN[4 * BP];

// Readers of the staged box.  Only faces whose whole stencil lies in the box
// use them (every interior face except the periodic wrap-around faces).
struct LdsU {
    int i0, j0;
    __device__ __forceinline__ double operator()(int k, int j, int i, int q) const
    {
        return sU[(((k & 3) * 3 + q) * BJ + (j - j0 + 1)) * BI + (i - i0 + 1)];
    }
};

struct LdsN {
    int i0, j0;
    __device__ __forceinline__ double operator()(int k, int j, int i) const
    {
        return (double)sN[((k & 3) * BJ + (j - j0 + 1)) * BI + (i - i0 + 1)];
    }
};

// Per-thread offsets of the staged box, fixed over the k march: for each
// ucat value a thread stages, its offset within a global plane (-1: outside
// the grid) and within an LDS plane; likewise for nvert.
struct StageMap {
    int gu[NU], lu[NU], gn[NN];
};

__device__ __forceinline__ StageMap stage_map(const Dims &d, int i0, int j0)
{
    StageMap m;
#pragma unroll
    for (int n = 0; n < NU; n++) {
        const int e = threadIdx.x + n * NTB;
        const int r = e / (3 * BI), c = e - r * (3 * BI);
        const int p = c / 3, q = c - 3 * p;
        const int gi = i0 - 1 + p, gj = j0 - 1 + r;
        const bool ok = e < 3 * BP && gi < d.mx && gj < d.my;
        m.gu[n] = ok ? 3 * (gj * d.mx + i0 - 1) + c : -1;
        m.lu[n] = e < 3 * BP ? (q * BJ + r) * BI + p : -1;
    }
#pragma unroll
    for (int n = 0; n < NN; n++) {
        const int e = threadIdx.x + n * NTB;
        const int r = e / BI, c = e - r * BI;
        const int gi = i0 - 1 + c, gj = j0 - 1 + r;
        m.gn[n] = e < BP && gi < d.mx && gj < d.my ? gj * d.mx + gi : -1;
    }
    return m;
}

// global loads of plane k of the staged box into registers
__device__ __forceinline__ void fetch_plane(const Dims &d, const double *__restrict__ ucat,
                                            const double *__restrict__ nvert, const StageMap &m, int k,
                                            double (&ru)[NU], float (&rn)[NN])
{
    const bool kin = k >= 0 && k < d.mz;
    const unsigned pl = (unsigned)k * (unsigned)d.mx * (unsigned)d.my;
#pragma unroll
    for (int n = 0; n < NU; n++) ru[n] = kin && m.gu[n] >= 0 ? ld32(ucat, 3 * pl + (unsigned)m.gu[n]) : 0.0;
#pragma unroll
    for (int n = 0; n < NN; n++) rn[n] = kin && m.gn[n] >= 0 ? (float)ld32(nvert, pl + (unsigned)m.gn[n]) : 0.0f;
}

__device__ __forceinline__ void store_plane(int k, const StageMap &m, const double (&ru)[NU], const float (&rn)[NN])
{
    const int slot = k & 3;
#pragma unroll
    for (int n = 0; n < NU; n++)
        if (m.lu[n] >= 0) sU[slot * 3 * BP + m.lu[n]] = ru[n];
#pragma unroll
    for (int n = 0; n < NN; n++) {
        const int e = threadIdx.x + n * NTB;
        if (e < BP) sN[slot * BP + e] = rn[n];
    }
}

}  // namespace

// The three face fluxes of point (k,j,i) from the staged planes, written to
// the face array.  The periodic wrap-around faces are left to
// k_ff2_faces_bnd.  PLAIN: no cell within reach of the tile's faces of this
// plane is marked, so every face except the last of its direction takes the
// mask-free specialisation of face_flux_t (bit-identical results).
template <int DIR, bool WENO, bool PLAIN>
__device__ __forceinline__ void face_dir(const Dims &d, const Geom &g, const double *__restrict__ ucont,
                                         const double *__restrict__ nu_t, double *__restrict__ face, int i0, int j0,
                                         int i, int j, int k)
{
    const LdsU U{i0, j0};
    const long long P = d.P;
    const unsigned t = IX32(d, k, j, i);
    const int n = DIR == 0 ? i : DIR == 1 ? j : k;
    const int m = DIR == 0 ? d.mx : DIR == 1 ? d.my : d.mz;
    double D[3], V[3];
    if (PLAIN && n != m - 2)
        face_flux_t<DIR, double, LdsU, UConPlain, ZeroN, StoredFaceMetric, WENO, unsigned, true>(
            d, g, ZeroN{}, UConPlain{ucont, DIR}, U, nu_t, i, j, k, D, V);
    else
        face_flux_t<DIR, double, LdsU, UConPlain, LdsN, StoredFaceMetric, WENO, unsigned>(
            d, g, LdsN{i0, j0}, UConPlain{ucont, DIR}, U, nu_t, i, j, k, D, V);
    // slot s of the face array starts at face + s * P (a uniform pointer)
#pragma unroll
    for (int q = 0; q < 3; q++) {
        st32(face + (DIR * 3 + q) * P, t, D[q]);
        st32(face + ((3 + DIR) * 3 + q) * P, t, V[q]);
    }
}

template <bool WENO, bool PLAIN>
__device__ __forceinline__ void faces_at(const Dims &d, const Geom &g, const double *__restrict__ ucont,
                                         const double *__restrict__ nu_t, double *__restrict__ face, int i0, int j0,
                                         int i, int j, int k)
{
    if (!(d.ip && i == d.mx - 2)) face_dir<0, WENO, PLAIN>(d, g, ucont, nu_t, face, i0, j0, i, j, k);
    // one face at a time: keeps the live registers of a single face
    __builtin_amdgcn_sched_barrier(0);
    if (!(d.jp && j == d.my - 2)) face_dir<1, WENO, PLAIN>(d, g, ucont, nu_t, face, i0, j0, i, j, k);
    __builtin_amdgcn_sched_barrier(0);
    if (!(d.kp && k == d.mz - 2)) face_dir<2, WENO, PLAIN>(d, g, ucont, nu_t, face, i0, j0, i, j, k);
}

// Plane flags of the face kernel: flag[(k * gy + by) * gx + bx] = 1 when no
// cell of the tile's staged box on planes k-1..k+2 (one-cell i/j halo, two on
// the high side, widened by one more each way) is marked.
// One wavefront a flag (grid gx * gy * mz, 64 threads): the lanes take the
// rows of the box in turn, a lane a point.
__global__ void __launch_bounds__(64) k_ff2_plain_flags(Dims d, const double *__restrict__ nvert, int gx, int gy,
                                                        unsigned char *__restrict__ flag)
{
    const long long t = blockIdx.x;
    const int bx = (int)(t % gx), by = (int)((t / gx) % gy), k = (int)(t / ((long long)gx * gy));
    const int i0 = 1 + bx * TI, j0 = 1 + by * TJ;
    const int ia = max(i0 - 3, 0), ib = min(i0 + TI + 2, d.mx - 1);
    const int ja = max(j0 - 3, 0), jb = min(j0 + TJ + 2, d.my - 1);
    const int ka = max(k - 1, 0), kb = min(k + 2, d.mz - 1);
    bool marked = false;
    for (int kk = ka; kk <= kb; kk++)
        for (int jj = ja; jj <= jb; jj++)
            for (int ii = ia + (int)threadIdx.x; ii <= ib; ii += 64) marked |= nvert[IX(d, kk, jj, ii)] != 0.0;
    const bool any = __ballot(marked) != 0;
    if (threadIdx.x == 0) flag[t] = (unsigned char)!any;
}

// Face fluxes of all three directions at every interior point
// (1..mx-2, 1..my-2, k in [kb, ke)).  grid: (ceil((mx-2)/64), ceil((my-2)/TJ), k chunks)
template <bool WENO>
__global__ void __launch_bounds__(NTB) __attribute__((amdgpu_waves_per_eu(VFS_FACES_WAVES, VFS_FACES_WAVES)))
k_ff2_faces_lds(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                const double *__restrict__ ucat, const double *__restrict__ nu_t, double *__restrict__ face,
                int kchunk, const unsigned char *__restrict__ plain)
{
    const int tx = threadIdx.x % TI, ty = (threadIdx.x / TI) % TJ;
    // logical tile of this workgroup: XCD x takes a contiguous range (x fastest)
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned lb = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
    const int bx = lb % gridDim.x, by = (lb / gridDim.x) % gridDim.y, bz = lb / (gridDim.x * gridDim.y);
    const int i0 = 1 + bx * TI, j0 = 1 + by * TJ;
    const int kb = 1 + bz * kchunk;
    const int ke = min(kb + kchunk, d.mz - 1);
    const int i = i0 + tx, j = j0 + ty;
    const bool active = i <= d.mx - 2 && j <= d.my - 2;
    const unsigned fstride = gridDim.x * gridDim.y, fbase = bx + gridDim.x * by;

    double ru[NU];
    float rn[NN];
    const StageMap sm = stage_map(d, i0, j0);
    for (int k = kb - 1; k <= kb + 2; k++) {
        fetch_plane(d, ucat, nvert, sm, k, ru, rn);
        store_plane(k, sm, ru, rn);
    }
    __syncthreads();
    for (int k = kb; k < ke; k++) {
        // prefetch plane k+3 while plane k is computed
        fetch_plane(d, ucat, nvert, sm, k + 3, ru, rn);
        if (active) {
            if (plain[fbase + fstride * (unsigned)k]) faces_at<WENO, true>(d, g, ucont, nu_t, face, i0, j0, i, j, k);
            else faces_at<WENO, false>(d, g, ucont, nu_t, face, i0, j0, i, j, k);
        }
        __syncthreads();   // every read of plane k-1 is done
        store_plane(k + 3, sm, ru, rn);
        __syncthreads();
    }
}

// Face points the staged kernel does not compute: the low boundary planes
// (i-faces at i = 0, j-faces at j = 0, k-faces at k = 0) and, in periodic
// directions, the last interior faces (i = mx-2, j = my-2, k = mz-2), whose
// 4th-order stencil wraps around to cell 2 (momentum.c:699-705).
// Segment s of the launch covers face direction s % 3 on plane 0 (s < 3) or
// on plane m-2 (s >= 3).
__global__ void __launch_bounds__(256) k_ff2_faces_bnd(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                                const double *__restrict__ ucat, const double *__restrict__ nu_t,
                                double *__restrict__ face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n3[3] = {(long long)(d.my - 2) * (d.mz - 2), (long long)(d.mx - 2) * (d.mz - 2),
                             (long long)(d.mx - 2) * (d.my - 2)};
    const int per[3] = {d.ip, d.jp, d.kp};
    const long long P = d.P;
    for (int s = 0; s < 6; s++) {
        const int dir = s % 3;
        if (s >= 3 && !per[dir]) continue;
        if (t >= n3[dir]) { t -= n3[dir]; continue; }
        int i, j, k;
        if (dir == 0) {
            i = s < 3 ? 0 : d.mx - 2; j = 1 + (int)(t % (d.my - 2)); k = 1 + (int)(t / (d.my - 2));
        } else if (dir == 1) {
            j = s < 3 ? 0 : d.my - 2; i = 1 + (int)(t % (d.mx - 2)); k = 1 + (int)(t / (d.mx - 2));
        } else {
            k = s < 3 ? 0 : d.mz - 2; i = 1 + (int)(t % (d.mx - 2)); j = 1 + (int)(t / (d.mx - 2));
        }
        const long long o = IX(d, k, j, i);
        double D[3], V[3], A[3];   // A: -skew only, which takes the passes
        if (dir == 0) face_flux<0>(d, g, nvert, ucont, ucat, nu_t, i, j, k, D, V, A);
        else if (dir == 1) face_flux<1>(d, g, nvert, ucont, ucat, nu_t, i, j, k, D, V, A);
        else face_flux<2>(d, g, nvert, ucont, ucat, nu_t, i, j, k, D, V, A);
        for (int q = 0; q < 3; q++) { face[(dir * 3 + q) * P + o] = D[q]; face[((3 + dir) * 3 + q) * P + o] = V[q]; }
        return;
    }
}

// Cell divergence (momentum.c:1506-1636), k-marching, with every load
// address known before any value arrives.  Along each direction the faces a
// cell can read are the four at cyclic offsets -2, -1, 0, +1 (the far faces
// of the 4th-order correction, the upwind face, the own face; across a
// periodic seam face f stands for face 1 + (f - 1) mod (m - 2), the
// reference's ghost copies): the 4th-order pair takes offsets (+1, -2),
// the second-order fallback (0, -1).  All of them, and the seven nvert
// values that decide, are loaded at the top of a step; the k faces move
// through a register ring.  Same operands in the same order as cell_fp.
constexpr int DI = 64, DJ = 4, DT = DI * DJ;

// face index f of a direction with m points, as stored: periodic seams wrap
// cyclically, other out-of-range indices are clamped (never selected)
__device__ __forceinline__ int face_cyc(int f, int m, int per)
{
    if (per) {
        const int p = m - 2, r = (f - 1) % p;
        return 1 + (r < 0 ? r + p : r);
    }
    return f < 0 ? 0 : (f > m - 1 ? m - 1 : f);
}

// the 4th-order correction keeps its default faces at cell n (momentum.c:1530-1580)
__device__ __forceinline__ bool far_faces(int n, int m, int per, double nsum)
{
    if (!per && (n == 1 || n == 2 || n == m - 3 || n == m - 2)) return false;
    return !(nsum > 0.1);
}

__global__ void __launch_bounds__(DT) k_ff2_cells_km(Dims d, const double *__restrict__ nvert,
                                                    const double *__restrict__ face, double *__restrict__ fp,
                                                    int kchunk)
{
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned lb = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
    const int bx = lb % gridDim.x, by = (lb / gridDim.x) % gridDim.y, bz = lb / (gridDim.x * gridDim.y);
    const int i = 1 + bx * DI + (int)(threadIdx.x % DI), j = 1 + by * DJ + (int)(threadIdx.x / DI);
    if (i > d.mx - 2 || j > d.my - 2) return;
    const int kb = 1 + bz * kchunk, ke = min(kb + kchunk, d.mz - 1);
    const long long P = d.P;
    const unsigned sk = (unsigned)d.mx * (unsigned)d.my, sy = (unsigned)d.mx;
    // i and j candidate faces (offsets -2, -1, 0, +1) within a plane
    const unsigned ia = IX32(d, 0, j, face_cyc(i - 2, d.mx, d.ip)), ib = IX32(d, 0, j, face_cyc(i - 1, d.mx, d.ip)),
                   ic = IX32(d, 0, j, i), id = IX32(d, 0, j, face_cyc(i + 1, d.mx, d.ip));
    const unsigned ja = IX32(d, 0, face_cyc(j - 2, d.my, d.jp), i), jb = IX32(d, 0, face_cyc(j - 1, d.my, d.jp), i),
                   jd = IX32(d, 0, face_cyc(j + 1, d.my, d.jp), i);
    const bool fo = !(d.second_order || d.inviscid);
    const double *D0 = face, *D1 = face + 3 * P, *D2 = face + 6 * P;
    const double *V0 = face + 9 * P, *V1 = face + 12 * P, *V2 = face + 15 * P;
    // ring: convective k-flux at faces k-2..k+1, viscous at k-1..k (cyclic indices)
    double c0[3], c1[3], c2[3], c3[3], v0[3], v1[3];
    {
        const unsigned p0 = face_cyc(kb - 2, d.mz, d.kp) * sk, p1 = face_cyc(kb - 1, d.mz, d.kp) * sk,
                       p2 = (unsigned)kb * sk, p3 = face_cyc(kb + 1, d.mz, d.kp) * sk;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            c0[q] = ld32(D2 + q * P, ic + p0); c1[q] = ld32(D2 + q * P, ic + p1);
            c2[q] = ld32(D2 + q * P, ic + p2); c3[q] = ld32(D2 + q * P, ic + p3);
            v0[q] = ld32(V2 + q * P, ic + p1); v1[q] = ld32(V2 + q * P, ic + p2);
        }
    }
    for (int k = kb; k < ke; k++) {
        const unsigned pk = (unsigned)k * sk;
        // every load of the step first
        const double nc = ld32(nvert, ic + pk), nim = ld32(nvert, ic + pk - 1), nip = ld32(nvert, ic + pk + 1),
                     njm = ld32(nvert, ic + pk - sy), njp = ld32(nvert, ic + pk + sy),
                     nkm = ld32(nvert, ic + pk - sk), nkp = ld32(nvert, ic + pk + sk);
        double ai[3][4], aj[3][4], vi[3][2], vj[3][2], nc3[3], nv3[3];
        const unsigned pn = face_cyc(k + 2, d.mz, d.kp) * sk, pv = face_cyc(k + 1, d.mz, d.kp) * sk;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            ai[q][0] = ld32(D0 + q * P, ia + pk); ai[q][1] = ld32(D0 + q * P, ib + pk);
            ai[q][2] = ld32(D0 + q * P, ic + pk); ai[q][3] = ld32(D0 + q * P, id + pk);
            aj[q][0] = ld32(D1 + q * P, ja + pk); aj[q][1] = ld32(D1 + q * P, jb + pk);
            aj[q][2] = ld32(D1 + q * P, ic + pk); aj[q][3] = ld32(D1 + q * P, jd + pk);
            vi[q][0] = ld32(V0 + q * P, ib + pk); vi[q][1] = ld32(V0 + q * P, ic + pk);
            vj[q][0] = ld32(V1 + q * P, jb + pk); vj[q][1] = ld32(V1 + q * P, ic + pk);
            nc3[q] = ld32(D2 + q * P, ic + pn); nv3[q] = ld32(V2 + q * P, ic + pv);
        }
        const bool fi = far_faces(i, d.mx, d.ip, nim + nc + nip);
        const bool fj = far_faces(j, d.my, d.jp, njm + nc + njp);
        const bool fk = far_faces(k, d.mz, d.kp, nkm + nc + nkp);
        const double den_i = fi ? 3. : 1., den_j = fj ? 3. : 1., den_k = fk ? 3. : 1.;
        double f[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double dv = (ai[q][2] - ai[q][1] + aj[q][2] - aj[q][1] + c2[q] - c1[q]);
            double r;
            if (!fo) {
                r = dv;
            } else {
                double d4 = 0;
                d4 += ((fi ? ai[q][3] : ai[q][2]) - (fi ? ai[q][0] : ai[q][1])) * (1. / den_i);
                d4 += ((fj ? aj[q][3] : aj[q][2]) - (fj ? aj[q][0] : aj[q][1])) * (1. / den_j);
                d4 += ((fk ? c3[q] : c2[q]) - (fk ? c0[q] : c1[q])) * (1. / den_k);
                r = 9. / 8. * dv + -1. / 8. * d4;
            }
            if (!d.inviscid) r += (vi[q][1] - vi[q][0] + vj[q][1] - vj[q][0] + v1[q] - v0[q]) / 1.0;
            f[q] = r;
        }
        const unsigned t = ic + pk;
        st32(fp, 3 * t, f[0]); st32(fp, 3 * t + 1, f[1]); st32(fp, 3 * t + 2, f[2]);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            c0[q] = c1[q]; c1[q] = c2[q]; c2[q] = c3[q]; c3[q] = nc3[q];
            v0[q] = v1[q]; v1[q] = nv3[q];
        }
    }
}

// the plane flags of the face kernel, recomputed when the mask has changed
static int ff2_plain_flags(vfs_ctx *c, int gx, int gy)
{
    const Dims d = c->d;
    const size_t n = (size_t)gx * gy * d.mz;
    if (c->ff2_flags && c->ff2_flags_version == c->nvert_version) return 0;
    if (!c->ff2_flags) VFS_CHECK_HIP(hipMalloc(&c->ff2_flags, n));
    hipLaunchKernelGGL(k_ff2_plain_flags, dim3((unsigned)n), dim3(64), 0, c->stream, d, c->nvert, gx, gy,
                       c->ff2_flags);
    VFS_CHECK_LAUNCH();
    c->ff2_flags_version = c->nvert_version;
    return 0;
}

int vfs_ff2_fused(vfs_ctx *c, const double *ucont, const double *ucat, const double *nu_t, double *rhs, double scale)
{
    // the staged kernels hold nvert as FP32 (exact for the integer masks) and
    // address the arrays with 32-bit byte offsets; -skew and -clark take the
    // reference-structured passes
    if (!c->nvert_f32 || !VFS_IX32_OK(c->d.P) || c->d.skew || c->d.clark)
        return vfs_ff2_passes(c, ucont, ucat, nu_t, rhs, scale);
    const Dims d = c->d;
    const int kchunk = VFS_FACES_KCHUNK;
    const dim3 grid((d.mx - 2 + TI - 1) / TI, (d.my - 2 + TJ - 1) / TJ, (d.mz - 2 + kchunk - 1) / kchunk);
    if (ff2_plain_flags(c, (int)grid.x, (int)grid.y)) return -1;
    if (d.weno || d.inviscid)
        hipLaunchKernelGGL(k_ff2_faces_lds<true>, grid, dim3(NTB), 0, c->stream, d, c->g, c->nvert, ucont, ucat, nu_t,
                           c->face, kchunk, c->ff2_flags);
    else
        hipLaunchKernelGGL(k_ff2_faces_lds<false>, grid, dim3(NTB), 0, c->stream, d, c->g, c->nvert, ucont, ucat, nu_t,
                           c->face, kchunk, c->ff2_flags);
    VFS_CHECK_LAUNCH();
    const long long n3[3] = {(long long)(d.my - 2) * (d.mz - 2), (long long)(d.mx - 2) * (d.mz - 2),
                             (long long)(d.mx - 2) * (d.my - 2)};
    const long long nb = n3[0] * (1 + d.ip) + n3[1] * (1 + d.jp) + n3[2] * (1 + d.kp);
    hipLaunchKernelGGL(k_ff2_faces_bnd, dim3(grid_for(nb, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, ucont,
                       ucat, nu_t, c->face);
    VFS_CHECK_LAUNCH();
    const int kc3 = VFS_CELLS_KCHUNK;
    const dim3 g3((d.mx - 2 + DI - 1) / DI, (d.my - 2 + DJ - 1) / DJ, (d.mz - 2 + kc3 - 1) / kc3);
    hipLaunchKernelGGL(k_ff2_cells_km, g3, dim3(DT), 0, c->stream, d, c->nvert, c->face, c->fp, kc3);
    VFS_CHECK_LAUNCH();
    return vfs_ff2_rhs_only(c, rhs, scale);
}
