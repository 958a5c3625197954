// Open k boundaries: an inlet on the k = 0 plane and an outlet on the
// k = mz-1 plane (bctype[4] = 5, bctype[5] = 4).  Calc_Inlet_Area
// (bcs.c:216-328), InflowFlux (bcs.c:329-706), OutflowFlux (bcs.c:801-870)
// and the outlet branch of FormBCS (bcs.c:2285-2418), without levelset or
// inflow fluctuations.  threshold = 0.1 (bcs.c:5).
//
// The face writes run one thread a face.  Each plane sum runs in one
// workgroup of 1024: thread t adds the faces t, t + 1024, ... of the plane
// (rows of mx-2 faces, j-major as the reference walks them), then the
// threads' partials meet in a fixed butterfly-and-wave order.  The order is
// the same on every run, so the sums are deterministic; they differ from the
// reference's single j, i loop only by the rounding of the association
// (about 1e-16 relative).
#include "vfs_internal.h"
#include "vfs_reduce.h"

namespace {

constexpr double THR = 0.1;

__device__ __forceinline__ double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// rhs.c:34-55 Contra2Cart_single
__device__ void c2c_single(const double *csi, const double *eta, const double *zet, const double *uc, double *u)
{
    const double det = csi[0] * (eta[1] * zet[2] - eta[2] * zet[1]) - csi[1] * (eta[0] * zet[2] - eta[2] * zet[0]) +
                       csi[2] * (eta[0] * zet[1] - eta[1] * zet[0]);
    const double det0 = uc[0] * (eta[1] * zet[2] - eta[2] * zet[1]) - uc[1] * (csi[1] * zet[2] - csi[2] * zet[1]) +
                        uc[2] * (csi[1] * eta[2] - csi[2] * eta[1]);
    const double det1 = -uc[0] * (eta[0] * zet[2] - eta[2] * zet[0]) + uc[1] * (csi[0] * zet[2] - csi[2] * zet[0]) -
                        uc[2] * (csi[0] * eta[2] - csi[2] * eta[0]);
    const double det2 = uc[0] * (eta[0] * zet[1] - eta[1] * zet[0]) - uc[1] * (csi[0] * zet[1] - csi[1] * zet[0]) +
                        uc[2] * (csi[0] * eta[1] - csi[1] * eta[0]);
    u[0] = det0 / det;
    u[1] = det1 / det;
    u[2] = det2 / det;
}

constexpr int PLANE_BS = 1024;

// bcs.c:262-274: fluid |zet| of each k plane, one workgroup a plane
__global__ void __launch_bounds__(PLANE_BS) k_plane_areas(Dims d, const double *__restrict__ zet,
                                                          const double *__restrict__ nvert, double *__restrict__ k_area)
{
    __shared__ double sh[PLANE_BS / 64];
    const int k = blockIdx.x;
    const long long nf = (long long)(d.mx - 2) * (d.my - 2);
    double a = 0;
    if (k >= 1 && k <= d.mz - 2)
        for (long long q = threadIdx.x; q < nf; q += PLANE_BS) {
            const int j = 1 + (int)(q / (d.mx - 2)), i = 1 + (int)(q % (d.mx - 2));
            const long long t = IX(d, k, j, i);
            if (nvert[t - (long long)d.mx * d.my] + nvert[t] < 0.1) a += norm3(&zet[3 * t]);
        }
    a = block_sum<PLANE_BS>(a, sh);
    if (threadIdx.x == 0) k_area[k] = a;
}

// bcs.c:275-310: k_area[0] = k_area[1], the mean over 1..mz-2 (in k order)
// and the fluid area of the inlet plane (one workgroup)
__global__ void __launch_bounds__(PLANE_BS) k_area_finish(Dims d, const double *__restrict__ zet,
                                                          const double *__restrict__ nvert,
                                                          double *__restrict__ k_area, double *__restrict__ out)
{
    __shared__ double sh[PLANE_BS / 64];
    const long long nf = (long long)(d.mx - 2) * (d.my - 2);
    double a = 0;
    for (long long q = threadIdx.x; q < nf; q += PLANE_BS) {
        const int j = 1 + (int)(q / (d.mx - 2)), i = 1 + (int)(q % (d.mx - 2));
        if (nvert[IX(d, 1, j, i)] < THR) a += norm3(&zet[3 * IX(d, 0, j, i)]);
    }
    a = block_sum<PLANE_BS>(a, sh);
    if (threadIdx.x == 0) {
        k_area[0] = k_area[1];
        double m = 0;
        for (int k = 1; k <= d.mz - 2; k++) m += k_area[k];
        out[0] = m / (double)(d.mz - 2);
        out[1] = a;
    }
}

// rhs2.c:1606-1619: the fluid flux through the k faces of each plane
// (fluid: nvert(k-1) + nvert(k) < 0.1), one workgroup a plane
__global__ void __launch_bounds__(PLANE_BS) k_kflux_planes(Dims d, const double *__restrict__ nvert,
                                                           const double *__restrict__ ucont, double *__restrict__ kf)
{
    __shared__ double sh[PLANE_BS / 64];
    const int k = blockIdx.x;
    const long long nf = (long long)(d.mx - 2) * (d.my - 2);
    double f = 0;
    if (k >= 1 && k <= d.mz - 2)
        for (long long q = threadIdx.x; q < nf; q += PLANE_BS) {
            const int j = 1 + (int)(q / (d.mx - 2)), i = 1 + (int)(q % (d.mx - 2));
            const long long t = IX(d, k, j, i);
            if (nvert[t - (long long)d.mx * d.my] + nvert[t] < 0.1) f += ucont[3 * t + 2];
        }
    f = block_sum<PLANE_BS>(f, sh);
    if (threadIdx.x == 0) kf[k] = f;
}

// the plane sums over k = 1..mz-2 in k order: Flux / (mz-2) (rhs2.c:1627)
// and the mean fluid area (bcs.c:277-278)
__global__ void k_kflux_finish(Dims d, const double *__restrict__ kf, const double *__restrict__ ka,
                               double *__restrict__ out)
{
    double f = 0, a = 0;
    for (int k = 1; k <= d.mz - 2; k++) {
        f += kf[k];
        a += ka[k];
    }
    out[0] = f / (d.mz - 2);
    out[1] = a / (double)(d.mz - 2);
}

// bcs.c:394-470: the inflow speed of the profiles
__device__ double inflow_speed(const vfs_inflow_options &o, double yc, const double *plane, long long ji)
{
    switch (o.inletprofile) {
    case 1:
    case 2: return o.inlet_flux < 0 ? 1. : o.inlet_flux / o.inlet_area;
    case 10: {
        const double delta = 0.45263, a = 5.99;
        if (yc >= delta) return 1.0;
        if (yc <= 0) return 0.0;
        return pow(yc / delta, 1. / a);
    }
    case 21: return 1.5 * 1.0 * yc * (2 - yc);
    default: return plane[ji];
    }
}

// bcs.c:383-595 (generic branch): the inlet face fluxes, the ghost velocity
// from them and ubcs, one thread a face of the k = 0 plane
__global__ void k_inflow_faces(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ coor,
                               vfs_inflow_options o, const double *__restrict__ plane, double *__restrict__ ucont,
                               double *__restrict__ ucat, double *__restrict__ ubcs)
{
    const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x), j = 1 + (int)blockIdx.y;
    if (i > d.mx - 2 || j > d.my - 2) return;
    const int k = 0;
    auto co = [&](int kk, int jj, int ii) { return coor[3 * IX(d, kk, jj, ii) + 1]; };
    const double yc = (co(k + 1, j, i) + co(k + 1, j - 1, i) + co(k + 1, j, i - 1) + co(k + 1, j - 1, i - 1)) * 0.25;
    const long long id = IX(d, k, j, i), up = IX(d, k + 1, j, i);
    const double area = norm3(&g.zet[3 * id]);
    const double uin = inflow_speed(o, yc, plane, (long long)j * d.mx + i);
    if (nvert[up] < THR) {
        double uc[3] = {0, 0, uin * area}, u[3];
        ucont[3 * id + 2] = uc[2];
        ucont[3 * id + 0] = 0;
        ucont[3 * id + 1] = 0;
        c2c_single(&g.csi[3 * id], &g.eta[3 * id], &g.zet[3 * id], uc, u);
        for (int q = 0; q < 3; q++) {
            const double v = -ucat[3 * up + q] + 2 * u[q];
            ucat[3 * id + q] = v;
            ubcs[3 * id + q] = v;
        }
    } else {
        ucat[3 * id + 2] = 0;
        ubcs[3 * id + 2] = 0;
        ucont[3 * id + 2] = 0;
    }
}

// the fluid flux through a k plane of faces (kf) whose fluid test reads
// plane kn: bcs.c:686-705 (inlet: kf = 0, kn = 1), 845-857 (outlet: kf = kn =
// mz-2); with area, also the fluid |zet| of plane kf (bcs.c:2299-2330).  One
// workgroup of PLANE_BS.
__global__ void __launch_bounds__(PLANE_BS) k_plane_flux(Dims d, const double *__restrict__ zet,
                                                         const double *__restrict__ nvert,
                                                         const double *__restrict__ ucont, int kf, int kn,
                                                         int with_area, double *__restrict__ out)
{
    __shared__ double sh[PLANE_BS / 64];
    const long long nf = (long long)(d.mx - 2) * (d.my - 2);
    double f = 0, a = 0;
    for (long long q = threadIdx.x; q < nf; q += PLANE_BS) {
        const int j = 1 + (int)(q / (d.mx - 2)), i = 1 + (int)(q % (d.mx - 2));
        if (nvert[IX(d, kn, j, i)] < THR) {
            const long long t = IX(d, kf, j, i);
            f += ucont[3 * t + 2];
            if (with_area) a += norm3(&zet[3 * t]);
        }
    }
    f = block_sum<PLANE_BS>(f, sh);
    a = block_sum<PLANE_BS>(a, sh);
    if (threadIdx.x == 0) {
        out[0] = f;
        out[1] = a;
    }
}

// bcs.c:2290-2297: ubcs of the outlet plane from the last cells
__global__ void k_outlet_ubcs(Dims d, const double *__restrict__ nvert, const double *__restrict__ ucat,
                              double *__restrict__ ubcs)
{
    const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x), j = 1 + (int)blockIdx.y;
    if (i > d.mx - 2 || j > d.my - 2) return;
    const long long id = IX(d, d.mz - 1, j, i), in = IX(d, d.mz - 2, j, i);
    ubcs[3 * id] = ucat[3 * in];
    ubcs[3 * id + 1] = ucat[3 * in + 1];
    ubcs[3 * id + 2] = nvert[in] < THR ? ucat[3 * in + 2] : ucat[3 * in + 2] * 0;
}

// bcs.c:2368-2398: the flux deficit spread over the outlet faces by area
__global__ void k_outlet_scale(Dims d, const double *__restrict__ zet, const double *__restrict__ nvert,
                               double flux_in, const double *__restrict__ sums, double *__restrict__ ucont)
{
    const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x), j = 1 + (int)blockIdx.y;
    if (i > d.mx - 2 || j > d.my - 2) return;
    const long long in = IX(d, d.mz - 2, j, i);
    const double A = norm3(&zet[3 * in]);
    if (nvert[in] < THR) ucont[3 * in + 2] += (flux_in - sums[0]) * A / sums[1];
}

int read_back(vfs_ctx *c, double *dst, int n)
{
    VFS_CHECK_HIP(hipMemcpyAsync(dst, c->red, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

dim3 plane_grid(const Dims &d) { return dim3((unsigned)((d.mx - 2 + 63) / 64), (unsigned)(d.my - 2)); }

}  // namespace

extern "C" int vfs_inlet_area(vfs_ctx *c, double *k_area, double *mean_k_area, double *inlet_area)
{
    const Dims d = c->d;
    VFS_REQUIRE(!d.kp, "open k boundaries need k_periodic = 0");
    VFS_REQUIRE(k_area && mean_k_area && inlet_area, "bad arguments");
    hipLaunchKernelGGL(k_plane_areas, dim3(d.mz), dim3(PLANE_BS), 0, c->stream, d, c->g.zet, c->nvert, k_area);
    hipLaunchKernelGGL(k_area_finish, dim3(1), dim3(PLANE_BS), 0, c->stream, d, c->g.zet, c->nvert, k_area, c->red);
    VFS_CHECK_LAUNCH();
    double h[2];
    if (read_back(c, h, 2)) return -1;
    *mean_k_area = h[0];
    *inlet_area = h[1];
    return 0;
}

extern "C" int vfs_calc_k_flux(vfs_ctx *c, const double *ucont, double *mean_k_flux, double *mean_k_area)
{
    const Dims d = c->d;
    VFS_REQUIRE(ucont, "bad arguments");
    double *ka = c->tmp3, *kf = c->tmp3 + d.mz;   // 2 mz doubles of scratch
    hipLaunchKernelGGL(k_plane_areas, dim3(d.mz), dim3(PLANE_BS), 0, c->stream, d, c->g.zet, c->nvert, ka);
    hipLaunchKernelGGL(k_kflux_planes, dim3(d.mz), dim3(PLANE_BS), 0, c->stream, d, c->nvert, ucont, kf);
    hipLaunchKernelGGL(k_kflux_finish, dim3(1), dim3(1), 0, c->stream, d, kf, ka, c->kflux);
    VFS_CHECK_LAUNCH();
    c->kflux_ready = 1;
    if (mean_k_flux || mean_k_area) {
        double h[2];
        VFS_CHECK_HIP(hipMemcpyAsync(h, c->kflux, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
        if (mean_k_flux) *mean_k_flux = h[0];
        if (mean_k_area) *mean_k_area = h[1];
    }
    return 0;
}

extern "C" int vfs_inflow_flux(vfs_ctx *c, const vfs_inflow_options *o, const double *coor, const double *uin_plane,
                               double *ucont, double *ucat, double *ubcs, double *flux_in)
{
    const Dims d = c->d;
    VFS_REQUIRE(!d.kp, "open k boundaries need k_periodic = 0");
    VFS_REQUIRE(o && coor && ucont && ucat && ubcs && flux_in, "bad arguments");
    const int pr = o->inletprofile;
    VFS_REQUIRE(pr == 1 || pr == 2 || pr == 10 || pr == 21 || (pr == 0 && uin_plane),
                "inletprofile must be 1, 2, 10, 21, or 0 with a uin plane");
    VFS_REQUIRE(!((pr == 1 || pr == 2) && o->inlet_flux >= 0) || o->inlet_area > 0,
                "a set inlet_flux needs the inlet area (vfs_inlet_area)");
    hipLaunchKernelGGL(k_inflow_faces, plane_grid(d), dim3(64), 0, c->stream, d, c->g, c->nvert, coor, *o, uin_plane,
                       ucont, ucat, ubcs);
    hipLaunchKernelGGL(k_plane_flux, dim3(1), dim3(PLANE_BS), 0, c->stream, d, c->g.zet, c->nvert, ucont, 0, 1, 0, c->red);
    VFS_CHECK_LAUNCH();
    double h[2];
    if (read_back(c, h, 2)) return -1;
    *flux_in = h[0];
    return 0;
}

extern "C" int vfs_outflow_flux(vfs_ctx *c, const double *ucont, double *flux_out)
{
    const Dims d = c->d;
    VFS_REQUIRE(!d.kp, "open k boundaries need k_periodic = 0");
    VFS_REQUIRE(ucont && flux_out, "bad arguments");
    hipLaunchKernelGGL(k_plane_flux, dim3(1), dim3(PLANE_BS), 0, c->stream, d, c->g.zet, c->nvert, ucont, d.mz - 2, d.mz - 2, 0,
                       c->red);
    VFS_CHECK_LAUNCH();
    double h[2];
    if (read_back(c, h, 2)) return -1;
    *flux_out = h[0];
    return 0;
}

extern "C" int vfs_outlet_correction(vfs_ctx *c, double flux_in, int scale, const double *ucat, double *ucont,
                                     double *ubcs, double *flux_out, double *area)
{
    const Dims d = c->d;
    VFS_REQUIRE(!d.kp, "open k boundaries need k_periodic = 0");
    VFS_REQUIRE(ucat && ucont && ubcs && flux_out && area, "bad arguments");
    hipLaunchKernelGGL(k_outlet_ubcs, plane_grid(d), dim3(64), 0, c->stream, d, c->nvert, ucat, ubcs);
    hipLaunchKernelGGL(k_plane_flux, dim3(1), dim3(PLANE_BS), 0, c->stream, d, c->g.zet, c->nvert, ucont, d.mz - 2, d.mz - 2, 1,
                       c->red);
    if (scale)
        hipLaunchKernelGGL(k_outlet_scale, plane_grid(d), dim3(64), 0, c->stream, d, c->g.zet, c->nvert, flux_in, c->red,
                           ucont);
    VFS_CHECK_LAUNCH();
    double h[2];
    if (read_back(c, h, 2)) return -1;
    *flux_out = h[0];
    *area = h[1];
    return 0;
}
