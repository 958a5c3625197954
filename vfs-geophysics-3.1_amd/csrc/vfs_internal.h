// Internal declarations shared by the HIP translation units of libvfsgpu.
//
// Data layout in HBM: every field is a flat array on the reference's DMDA
// single-rank layout (mx*my*mz points, i fastest), vectors as {x,y,z} AoS.
// Kernels are launched on the context's stream; nothing here allocates per
// call except the one-time scratch owned by the context.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string>

#include "vfsgpu.h"

#define VFS_HOST_DEVICE __host__ __device__ __forceinline__

// Grid description passed by value to every kernel.
struct Dims {
    int mx, my, mz;
    int ip, jp, kp;          // i/j/k_periodic
    int bc[6];               // bctype
    int immersed, les, second_order, dpdz_set, weno, inviscid, skew, clark;
    double ren, dt, st, mpg;
    long long P;             // mx*my*mz
};

VFS_HOST_DEVICE long long IX(const Dims &d, int k, int j, int i)
{
    return ((long long)k * d.my + j) * (long long)d.mx + i;
}

// 32-bit indexing for the hot kernels: the point index in 32 bits, and array
// elements read by (uniform base pointer) + 32-bit byte offset, which the
// loads take as a scalar base and a vector offset with no 64-bit address
// arithmetic.  Valid while every byte offset fits 32 bits (24 * P < 2^32,
// checked on the host before such a kernel is launched).
VFS_HOST_DEVICE unsigned IX32(const Dims &d, int k, int j, int i)
{
    return ((unsigned)k * (unsigned)d.my + (unsigned)j) * (unsigned)d.mx + (unsigned)i;
}
template <typename T>
__device__ __forceinline__ T ld32(const T *p, unsigned e)
{
    return *(const T *)((const char *)p + e * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T *p, unsigned e, T v)
{
    *(T *)((char *)p + e * (unsigned)sizeof(T)) = v;
}
// element e of p with either index type
template <typename T> __device__ __forceinline__ T ldx(const T *p, long long e) { return p[e]; }
template <typename T> __device__ __forceinline__ T ldx(const T *p, unsigned e) { return ld32(p, e); }
template <typename Idx> __device__ __forceinline__ Idx ixt(const Dims &d, int k, int j, int i);
template <> __device__ __forceinline__ long long ixt<long long>(const Dims &d, int k, int j, int i) { return IX(d, k, j, i); }
template <> __device__ __forceinline__ unsigned ixt<unsigned>(const Dims &d, int k, int j, int i) { return IX32(d, k, j, i); }
#define VFS_IX32_OK(P) (24ull * (unsigned long long)(P) < (1ull << 32))

// slots of the face array: div1..3, visc1..3 and, under -skew, adv1..3 (3 each)
#define VFS_FACE_SLOTS(d) ((d).skew ? 27 : 18)

// Periodic ghost source of point (k,j,i): the reference's `flag/a/b/c` idiom
// (e.g. momentum.c:642-669).  Sources are never flagged themselves.
VFS_HOST_DEVICE bool per_src(const Dims &d, int i, int j, int k, int &a, int &b, int &c)
{
    bool flag = false;
    a = i; b = j; c = k;
    if (d.ip && i == 0) a = d.mx - 2, flag = true;
    else if (d.ip && i == d.mx - 1) a = 1, flag = true;
    if (d.jp && j == 0) b = d.my - 2, flag = true;
    else if (d.jp && j == d.my - 1) b = 1, flag = true;
    if (d.kp && k == 0) c = d.mz - 2, flag = true;
    else if (d.kp && k == d.mz - 1) c = 1, flag = true;
    return flag;
}

// Points of the ghost shell (any index on a boundary plane), enumerated as
// the two k planes, then the two j planes without k ghosts, then the two i
// planes without j/k ghosts.
VFS_HOST_DEVICE long long shell_count(const Dims &d)
{
    long long A = (long long)d.mx * d.my, B = (long long)d.mx * (d.mz - 2),
              C = (long long)(d.my - 2) * (d.mz - 2);
    return 2 * (A + B + C);
}

VFS_HOST_DEVICE bool shell_point(const Dims &d, long long t, int &k, int &j, int &i)
{
    long long A = (long long)d.mx * d.my;
    if (t < 2 * A) {
        k = t < A ? 0 : d.mz - 1;
        long long r = t % A;
        j = (int)(r / d.mx); i = (int)(r % d.mx);
        return true;
    }
    t -= 2 * A;
    long long B = (long long)d.mx * (d.mz - 2);
    if (t < 2 * B) {
        j = t < B ? 0 : d.my - 1;
        long long r = t % B;
        k = 1 + (int)(r / d.mx); i = (int)(r % d.mx);
        return true;
    }
    t -= 2 * B;
    long long C = (long long)(d.my - 2) * (d.mz - 2);
    if (t < 2 * C) {
        i = t < C ? 0 : d.mx - 1;
        long long r = t % C;
        k = 1 + (int)(r / (d.my - 2)); j = 1 + (int)(r % (d.my - 2));
        return true;
    }
    return false;
}

struct Geom {
    double *csi, *eta, *zet, *aj;
    double *icsi, *ieta, *izet, *iaj;
    double *jcsi, *jeta, *jzet, *jaj;
    double *kcsi, *keta, *kzet, *kaj;
    double *cent;
};

struct Poisson;   // defined in vfs_solver.hip

struct vfs_ctx {
    vfs_config cfg;
    Dims d;
    hipStream_t stream;
    int device;
    Geom g;
    double *geom_block;      // one allocation holding every metric array
    double *nvert;
    int nvert_f32;           // every nvert value is exactly representable in FP32
    // scratch (device)
    double *face;            // 18*P: div1..3 and visc1..3 (vector each); 27*P with adv1..3 under -skew
    double *fp;              // 3*P
    double *snap;            // 3*P: Contra2Cart boundary snapshot / SNES copies
    double *tmp3;            // 3*P general
    double *red;             // reduction partials + result scalars
    double *ad;              // 12*P: primal and adjoint state of the JVP/VJP (allocated on first use)
    double *vjp_ws;          // 26*P: the face adjoints of one direction in the VJP (allocated on first use)
    void *nk;                // Newton-Krylov workspace (allocated on first use)
    double *wf_table;        // wall-function y+ integral table (bctype -1 only)
    double *ustar;           // P: friction velocity of the wall-function cells
    double *dsm;             // 19*P: dynamic Smagorinsky scratch (allocated on first use)
    const double *body_force;   // F_eul added to the implicit residual (caller's array, may be null)
    double *kflux;           // 2 doubles: mean_k_flux (Calc_k_Flux) and mean_k_area (Calc_Inlet_Area)
    int kflux_ready;         // vfs_calc_k_flux has run
    double inlet_flux;       // -flux (main.c:198, 2067): the target k flux of the forcing, default -1
    int nvert_version;       // bumped by every entry point that writes nvert
    unsigned char *ff2_flags;   // mask-free planes of the face kernel's tiles (allocated on first use)
    int ff2_flags_version;   // nvert_version the flags were computed for
    int red_blocks;
    Poisson *pois;
};

// the 3x3x3 dilation and 3 / 1 relabelling of an inside-marked mask (vfs_metrics.hip)
int vfs_mask_dilate_relabel(vfs_ctx *c);

// the device address of the static scratch of vfs_ib_search_noalloc,
// recorded once per device (vfs_ibm.hip)
int vfs_ib_scratch_setup(int device);

// wall-function walls (vfs_wall.hip)
bool vfs_has_wallfn(const Dims &d);
int vfs_wall_setup(vfs_ctx *c);
int vfs_wall_mask(vfs_ctx *c);
int vfs_c2c_wall(vfs_ctx *c, const double *L, double *ucat);
int vfs_ibbc_wall(vfs_ctx *c, double *ucont);

// error handling
void vfs_set_error(const std::string &msg);
#define VFS_CHECK_HIP(expr)                                                      \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            vfs_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
            return -1;                                                           \
        }                                                                        \
    } while (0)
#define VFS_CHECK_LAUNCH()                                                       \
    do {                                                                         \
        hipError_t _e = hipGetLastError();                                       \
        if (_e != hipSuccess) {                                                  \
            vfs_set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
            return -1;                                                           \
        }                                                                        \
    } while (0)
#define VFS_REQUIRE(cond, msg)                                                   \
    do {                                                                         \
        if (!(cond)) { vfs_set_error(msg); return -2; }                          \
    } while (0)

static inline unsigned grid_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// Workgroups are dispatched to the 8 XCDs of the MI355X round-robin by
// linear workgroup id, and each XCD has its own L2.  xcd_block maps the
// hardware id to a logical id so that XCD x runs one contiguous range of
// logical blocks (a bijection for any nb): neighbouring rows and planes of a
// stencil then meet in the same L2.
#define VFS_NUM_XCD 8
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
#ifdef VFS_NO_XCD_REMAP
    (void)nb;
    return b;
#else
    const unsigned q = nb / VFS_NUM_XCD, rem = nb % VFS_NUM_XCD;
    const unsigned x = b % VFS_NUM_XCD, r = b / VFS_NUM_XCD;
    return x * q + (x < rem ? x : rem) + r;
#endif
}

// shared launch helpers (vfs_core.hip)
int vfs_periodic_vec(vfs_ctx *c, double *v);      // periodic ghost copy, vector field
int vfs_periodic_sca(vfs_ctx *c, double *v);      // periodic ghost copy, scalar field
// deterministic reductions over n doubles produced by a per-block kernel
int vfs_reduce_finish(vfs_ctx *c, int nblocks, int nvals, int op /*0 sum, 1 max*/, double *dev_out);
int vfs_reduce_finish_slots(vfs_ctx *c, int first_slot, int nblocks, int nvals, int op, double *dev_out);
int vfs_read_scalars(vfs_ctx *c, const double *dev, double *host, int n);
int vfs_poisson_destroy(vfs_ctx *c);
void vfs_implicit_destroy(vfs_ctx *c);

// fixed sizes of the reduction buffer
#define VFS_RED_BLOCKS 1024
#define VFS_RED_SLOTS 64
