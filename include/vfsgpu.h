/*
 * vfsgpu.h -- C ABI of the MI355X (gfx950) implementation of the
 * VFS-Geophysics 3.1 fractional-step hot path: the curvilinear momentum
 * right-hand side and the pressure projection.
 *
 * Conventions (shared by every entry point):
 *  - A context (vfs_ctx) stands for the reference's UserCtx: grid sizes,
 *    flags, metrics and the nvert mask live in it, on the device.
 *  - Field arguments are DEVICE pointers in the reference's single-rank DMDA
 *    layout: sizes mx = IM+1, my = JM+1, mz = KM+1 (IM = node count), scalar
 *    s[(k*my + j)*mx + i], vector (Cmpnts) v[3*((k*my + j)*mx + i) + c] with
 *    c = x,y,z.  A host wrapper can hand DAVecGetArray storage straight
 *    through after copying it to the device.
 *  - Return value 0 = success, < 0 = error (vfs_last_error() explains); no
 *    entry point calls exit().  Work is enqueued on the context's stream;
 *    entry points that return scalars synchronize that stream.
 *  - All arithmetic of the drop-in path is IEEE FP64; the FP32 forms are
 *    explicit options (precision argument of vfs_poisson_solve).
 */
#ifndef VFSGPU_H
#define VFSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vfs_ctx vfs_ctx;

/* Run-time flags of the reference that the hot path reads (main.c:1602-2110,
 * init.c:52-69, bcs.dat via init.c:561-566).  bctype: 1 = no-slip wall,
 * 7 = periodic (the matching *_periodic flag must be set), -1 = wall
 * function (Cabot wall model, wallfunction.c) on an i or j wall: without an
 * immersed body its first cells become IB nodes (momentum.c:2014-2039). */
typedef struct {
    int mx, my, mz;
    int i_periodic, j_periodic, k_periodic;
    int bctype[6];
    int immersed;        /* -imm: 0 no IB, 1 CURVIB masks              */
    int les;             /* -les 1: constant-coefficient Smagorinsky    */
    int second_order;    /* -second_order                              */
    int dpdz_set;        /* k-periodic forcing from mean_pressure_gradient; 0: the
                            flux forcing (vfs_calc_k_flux, vfs_set_inlet_flux) */
    int weno;            /* -weno 1: WENO3 convective face values (level.c:1734) */
    int inviscid;        /* -inv: WENO3 convection, 2nd-order divergence, no
                            viscous term (momentum.c:757, 1582, 1621) */
    int skew;            /* -skew: skew-symmetric convection, half the divergence
                            form plus half the advective form interpolated to the
                            cells (momentum.c:798-811, 1060-1071, 1298-1309,
                            1596-1608) */
    int clark;           /* -clark: the Clark gradient-model stress in the viscous
                            face fluxes (momentum.c:907-925, 1142-1160, 1385-1403) */
    double ren, dt, st;  /* Reynolds number, time step, Strouhal factor */
    double mean_pressure_gradient;
} vfs_config;

/* ---- context, memory, stream ------------------------------------------ */
/* Replaces the UserCtx set-up of MG_Initial (init.c:270-930). */
int vfs_ctx_create(const vfs_config *cfg, vfs_ctx **out);
int vfs_ctx_destroy(vfs_ctx *ctx);
const char *vfs_last_error(void);
int vfs_set_stream(vfs_ctx *ctx, void *hip_stream);
int vfs_synchronize(vfs_ctx *ctx);
int vfs_device_malloc(size_t bytes, void **ptr);   /* zero-filled */
int vfs_device_free(void *ptr);
int vfs_memcpy_htod(void *dst, const void *src, size_t bytes);
int vfs_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int vfs_memcpy_dtod(void *dst, const void *src, size_t bytes);

/* ---- geometry ---------------------------------------------------------- */
/* FormMetrics (metrics.c:4-1346, NEWMETRIC): cell and face metrics from node
 * coordinates coor (vector field, nodes i<mx-1, j<my-1, k<mz-1).  Returns -7
 * when an interior cell's Jacobian is not a positive finite number (a
 * folded or degenerate grid; the reference prints "Negative jacobian",
 * momentum.c:157); the metrics are still written. */
int vfs_build_metrics(vfs_ctx *ctx, const double *coor);
/* Copy one metric array out: which = 0..16 in the order csi, eta, zet, aj,
 * icsi, ieta, izet, iaj, jcsi, jeta, jzet, jaj, kcsi, keta, kzet, kaj, cent. */
int vfs_get_metric(vfs_ctx *ctx, int which, double *dst);
/* Replace the context's nvert mask (the reference's user->Nvert). */
int vfs_set_nvert(vfs_ctx *ctx, const double *nvert);
int vfs_get_nvert(vfs_ctx *ctx, double *dst);
/* IB mask of a sphere from its signed distance at cell centres with the
 * reference's 3x3x3 dilation and relabelling (ibm.c:224, 319-341, 508-516):
 * 3 solid, 1 IB node, 0 fluid.  Stored in the context and copied to nvert_out
 * when that is not NULL. */
int vfs_nvert_sphere(vfs_ctx *ctx, double cx, double cy, double cz, double r, double *nvert_out);

/* ---- momentum ---------------------------------------------------------- */
/* Contra2Cart_2 (rhs.c:57-741; called through Contra2Cart, rhs.c:27).
 * ucont: in/out (periodic ghosts refreshed), ucat: in/out (values at cells
 * with nvert >= 0.1 are inputs and are kept, solid cells are zeroed). */
int vfs_contra2cart(vfs_ctx *ctx, double *ucont, double *ucat);
/* The friction velocity user->lUstar that Contra2Cart_2 writes on the first
 * cells of wall-function walls (rhs.c:303-401; zero elsewhere).  dst: P
 * doubles on the device.  Only for contexts with a bctype -1 wall. */
int vfs_get_ustar(vfs_ctx *ctx, double *dst);
/* Compute_eddy_viscosity_LES (les.c:1135-1252) with constant Cs
 * (Compute_Smagorinsky_Constant_1, les.c:67-77: Cs = 0.01 for -les 1). */
int vfs_eddy_viscosity(vfs_ctx *ctx, const double *ucat, double cs, double *nu_t);
/* The same with a per-cell Cs field (user->lCs, les.c:1171-1203): cs, P
 * doubles on the device, as vfs_smagorinsky_constant writes it. */
int vfs_eddy_viscosity_cs(vfs_ctx *ctx, const double *ucat, const double *cs, double *nu_t);
/* The options of Compute_Smagorinsky_Constant_1 (main.c:202-255, 1721-1772,
 * 2066). */
typedef struct {
    int les;                /* 1: Cs = 0.01; 2: dynamic procedure; 3: dynamic with
                               temporal relaxation (les.c:777-782) */
    int i_homo_filter;      /* -i_homo_filter: average LM, MM along i ... */
    int j_homo_filter;      /* -j_homo_filter */
    int k_homo_filter;      /* -k_homo_filter (i and k together: i-k planes) */
    int testfilter_ik;      /* -testfilter_ik: Simpson filter in i-k planes */
    int reserved;           /* set to 0 */
    double max_cs;          /* -max_cs (the reference's default 0.5) */
} vfs_les_options;
/* Compute_Smagorinsky_Constant_1 (les.c:67-1133) into cs (P doubles on the
 * device, every point written; with -les 3 it is also read: the previous
 * step's Cs).  The dynamic procedure (les 2, 3) is the reference's without
 * the Clark term (-clark).  Cs = 0 while ti < 2 on a fresh start (tistart 0,
 * rstart 0).  ucat must carry its ghost values (after vfs_contra2cart).
 * Scratch of 19 doubles a point is allocated on the first dynamic call and
 * kept. */
int vfs_smagorinsky_constant(vfs_ctx *ctx, const double *ucat, const vfs_les_options *opt, int ti, int tistart,
                             int rstart, double *cs);
/* ---- actuator models (rotor_model.c) ------------------------------------
 * Element arrays are device arrays of n values in blocks: cent x|y|z (3n),
 * f_lagr x|y|z (3n), dA (n), box i_min|i_max|j_min|j_max|k_min|k_max (6n,
 * upper bounds exclusive, the reference's IBMNodes fields), dh x|y|z (3n).
 * coor: the node coordinates given to vfs_build_metrics. */
/* The element search of the rotor pre-processing (rotor_model.c:2590-2737):
 * the closest grid node of each element centre (every 20th node first, then
 * a +-21 window; the first minimum in i, j, k loop order), the grid widths
 * there, and the index box closest node +- itpwidth (the reference's
 * Itpwidth, (int)halfwidth_dfunc + 2 = 6) clamped to the interior. */
int vfs_rotor_elements(vfs_ctx *ctx, const double *coor, int n, const double *cent, int itpwidth, int *box,
                       double *dh);
/* Calc_F_eul (rotor_model.c:4792-5137): f_eul (3P, every point written) =
 * sum over elements, in order, of F_lagr dA times the discrete delta
 * function of each face centre's offset (covariant directions, scaled by the
 * widths), projected on csi / eta / zet; zero where the 3x3x3 nvert sum
 * exceeds 2.9.  deltafunc: 0 dfunc_2h, 6 dfunc_nh(halfwidth), 7
 * dfunc_exp(halfwidth), 8 dfunc_s4h, otherwise dfunc_sc4h (the reference's
 * default 10).  width_mode 0: widths from the cell metrics (default); 1:
 * dh_fixed (3 host doubles, -forcewidthfixed); 2: dA, 3: sqrt(dA)
 * (-forcedistrwidth_surfacescale).  Boxes must lie in the interior, as
 * vfs_rotor_elements makes them.  Synchronizes the context's stream. */
int vfs_calc_f_eul(vfs_ctx *ctx, const double *coor, int n, const double *cent, const double *f_lagr,
                   const double *dA, const int *box, int deltafunc, double halfwidth, int width_mode,
                   const double *dh_fixed, double *f_eul);
/* Attach F_eul (3P doubles on the device, kept by the caller) to
 * vfs_formfunction_snes and the Newton-Krylov residual, which add it last
 * (momentum.c:2288, nacelle_model / rotor_model); NULL detaches. */
int vfs_set_body_force(vfs_ctx *ctx, const double *f_eul);

/* ---- CURVIB velocity reconstruction (ibm.c) -----------------------------
 * A triangulated body as the reference's IBMNodes holds it: n_v vertices,
 * n_elmt elements with unit normals pointing into the fluid, the vertex
 * velocities (3*n_v, Cmpnts AoS; NULL = a body at rest).  Device arrays. */
typedef struct {
    int n_v, n_elmt;
    const double *x_bp, *y_bp, *z_bp;
    const int *nv1, *nv2, *nv3;
    const double *nf_x, *nf_y, *nf_z;
    const double *u;
} vfs_ib_surface;
/* One IB node and the geometry of its reconstruction, the fields of IBMInfo
 * (variables.h:126-172) that ibm_search_advanced fills. */
typedef struct {
    int i, j, k;                    /* the node (ni, nj, nk) */
    int cell;                       /* the nearest element */
    int imode;                      /* the patch triangle hit (0..7); -100: the boundary fallback */
    int i1, j1, k1, i2, j2, k2, i3, j3, k3;   /* the three interpolation cells */
    double cr1, cr2, cr3;           /* their weights */
    double cs1, cs2, cs3;           /* the nearest point's weights on the element */
    double d_s;                     /* node to surface */
    double d_i;                     /* node to the interception point */
    double pmin[3];                 /* the nearest point on the surface */
} vfs_ib_node;
/* The marking of ibm_search_advanced (ibm.c:110-360, 508-516) for one body:
 * every interior cell whose centre lies inside the body's box (widened by
 * 0.05) casts a ray towards a jittered point above the body
 * (randomdirection, ibm.c:52-84; point_cell_advanced, ibm.c:1068-1183)
 * against the elements binned on a 40^3 grid; an odd count of crossings
 * marks it inside, two crossings closer than 1e-8 cast the next ray.  Then
 * the 3x3x3 dilation and the relabelling: 3 solid, 1 IB node, as
 * vfs_nvert_sphere.  reset: the first body (nvert < 5 cleared, ibm.c:247);
 * later bodies add to the mask.  -3 if a cell finds no clear ray in 32
 * tries.  The context's mask is set; nvert_out (P doubles) may be NULL.
 * Synchronizes the stream. */
int vfs_nvert_body(vfs_ctx *ctx, const vfs_ib_surface *surf, int reset, double *nvert_out);
/* The same with -thin (ibm.c:281-317, point_cell_thin ibm.c:1186-1244): for
 * a body of zero thickness (a plate, a blade), a cell whose segments to its
 * +i, +j, +k and +ijk neighbours cut an element marks the cells of that
 * 2x2x2 box below 0.5 with 2 (IB nodes after the relabelling). */
int vfs_nvert_body_thin(vfs_ctx *ctx, const vfs_ib_surface *surf, int reset, int thin, double *nvert_out);
/* The ray jitter of randomdirection: the first rand() after srand(seed) of
 * the C library, / (RAND_MAX + 1) - 0.5 (host only; the caller's rand()
 * sequence is left alone). */
int vfs_ib_ray_jitter(int seed, double *s);
/* The node list of ibm_search_advanced (ibm.c:366-497) for a mask already
 * set (vfs_set_nvert / vfs_nvert_sphere): every interior cell with
 * (int)(nvert+0.5) == 1, except the planes of wall-function walls, gets the
 * nearest element and point (nearestcell1, ibm.c:1366-1511, with the
 * bounding spheres of ibm.c:1276-1364) and the interception of the normal
 * ray with the cell-centre patches (InterceptionPoint, ibm.c:1759-2250), or
 * the boundary fallback (ibm.c:413-493).  nodes: device array of cap
 * entries, written in the reference's list order (AddIBMNode prepends: the
 * last cell in k, j, i order comes first); *n_nodes (host) the count.  -3 if
 * a node has no element or no interception (the reference prints "IBM
 * Searching Fail!"), -4 if cap is too small.  Synchronizes the stream. */
int vfs_ib_search(vfs_ctx *ctx, const vfs_ib_surface *surf, vfs_ib_node *nodes, int cap, int *n_nodes);
/* vfs_ib_search without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (tight loops that move a body
 * every step, harnesses that forbid hipMalloc).  Same arguments, results
 * and error codes, plus -5 when the job needs more than the library's static
 * 8 MiB scratch on the device (32 bytes per element plus 4 bytes per 256
 * grid points).  Every call on a device shares that device's scratch, which
 * keeps the last call's contents: calls that overlap on one device race.
 * Synchronizes the stream. */
int vfs_ib_search_noalloc(vfs_ctx *ctx, const vfs_ib_surface *surf, vfs_ib_node *nodes, int cap, int *n_nodes);
/* ibm_interpolation_advanced (ibm.c:3294-3698) without averaging, levelset,
 * FSI or an IB wall model: ucat at each node from the interpolated velocity
 * along the normal (noslip, or freeslip with -inv, wallfunction.c:134-166;
 * zero when an interpolation cell is not fluid or the node is solid), p from
 * the interpolated pressure, ustar (a P field) from the noslip model; then
 * the element means count / ustar^2 / p (n_elmt each, may all be NULL), the
 * fluxes of the faces of every nvert == 1 cell from the new ucat, and
 * Contra2Cart.  ucont, ucat, p, ustar: in/out device fields.  Reads see the
 * fields as they were before the call.  nodes: as vfs_ib_search writes them;
 * an entry whose cells or element fall outside the grid or the body is
 * skipped. */
int vfs_ib_interpolate(vfs_ctx *ctx, const vfs_ib_surface *surf, const vfs_ib_node *nodes, int n, double *ucont,
                       double *ucat, double *p, double *ustar, int *count, double *shear, double *pressure);

/* ---- open k boundaries (bcs.c): an inlet on the k = 0 plane (bctype[4] =
 * 5) and an outlet on the k = mz-1 plane (bctype[5] = 4) ------------------
 * The routines need k_periodic = 0; they act on those planes whatever the
 * context's k bctypes say, so the caller runs them where the reference does
 * (InflowFlux / OutflowFlux / FormBCS before each solve, implicitsolver.c
 * 4286-4396).  The other entry points keep treating the k ends as walls.
 * Fluid faces: nvert < 0.1 (threshold, bcs.c:5).  Sums run in the
 * reference's j, i order.  Each call synchronizes the stream. */
typedef struct {
    int inletprofile;       /* -inlet: 1 or 2 uniform (1, or inlet_flux / inlet_area
                               when inlet_flux >= 0), 10 the power law of
                               bcs.c:432-437, 21 1.5 y (2 - y); 0: the caller's
                               uin_plane (a profile the reference reads from a file) */
    int reserved;           /* set to 0 */
    double inlet_flux;      /* -inlet_flux (the reference's default -1: not set) */
    double inlet_area;      /* from vfs_inlet_area */
} vfs_inflow_options;
/* Calc_Inlet_Area (bcs.c:216-328): k_area (mz doubles on the device: the
 * fluid |zet| of each k plane, k_area[0] = k_area[1]), its mean over
 * 1..mz-2 and the fluid area of the inlet plane (host doubles). */
int vfs_inlet_area(vfs_ctx *ctx, double *k_area, double *mean_k_area, double *inlet_area);
/* InflowFlux (bcs.c:329-706) without levelset or fluctuations: at each face
 * of the k = 0 plane whose cell above is fluid, U^zeta = uin |zet|, the
 * other components 0, the ghost velocity 2 u - ucat(k = 1) with u =
 * Contra2Cart_single of that flux, and ubcs = the ghost velocity; else the
 * z components 0.  *flux_in (host): the fluid inflow.  coor: the node
 * coordinates given to vfs_build_metrics; uin_plane: mx*my doubles
 * (j*mx + i) for inletprofile 0, else NULL.  ucont, ucat, ubcs: device
 * fields (3P). */
int vfs_inflow_flux(vfs_ctx *ctx, const vfs_inflow_options *opt, const double *coor, const double *uin_plane,
                    double *ucont, double *ucat, double *ubcs, double *flux_in);
/* OutflowFlux (bcs.c:801-870): the fluid flux through the k = mz-2 faces. */
int vfs_outflow_flux(vfs_ctx *ctx, const double *ucont, double *flux_out);
/* FormBCS's outlet branch (bcs.c:2285-2418, no levelset): ubcs of the
 * k = mz-1 plane from the last cells, the fluid flux and area of the
 * k = mz-2 faces (host), and with scale (outflow_scale) the deficit
 * flux_in - flux_out added to those faces in proportion to |zet|. */
int vfs_outlet_correction(vfs_ctx *ctx, double flux_in, int scale, const double *ucat, double *ucont, double *ubcs,
                          double *flux_out, double *area);

/* Options of ibm_interpolation_advanced beyond the defaults of
 * vfs_ib_interpolate. */
typedef struct {
    int moving;             /* movefsi / rotatefsi: the fluid-support guard is off (ibm.c:3448) */
    int wallfunction;       /* -wallfunction at ti > 0: 1, the node velocity from the rough-wall
                               law of wall_function_roughness_a (sediment_transport.c:18-96);
                               2, with -roughness on body 0 (ibm.c:3468-3471), the rough log law
                               u = u_tau (ln(y / roughness) / 0.41 + 8.5) of
                               wall_function_roughness_loglaw (wallfunction.c:71-95), which keeps
                               the tangential part only; the averages use law 1 for both */
    double ks;              /* law 1's roughness height (the reference passes 0.0001, ibm.c:3473) */
    double roughness;       /* -roughness (main.c:2073): law 2's roughness height */
} vfs_ib_options;
/* vfs_ib_interpolate with options (ibm.c:3448-3478).  ks is read by law 1
 * only.  The reference takes its wall laws from ti = 1 on; at ti = 0 a
 * -wallfunction run takes noslip, even with -inv (ibm.c:3453-3477).  Here
 * wallfunction = 0 gives noslip on a viscous context and freeslip on an
 * inviscid one, so that one step of an inviscid -wallfunction run has no
 * counterpart. */
int vfs_ib_interpolate_opt(vfs_ctx *ctx, const vfs_ib_surface *surf, const vfs_ib_node *nodes, int n,
                           const vfs_ib_options *opt, double *ucont, double *ucat, double *p, double *ustar,
                           int *count, double *shear, double *pressure);

/* The averaging statistics of ibm_interpolation_advanced (ibm.c:3481-3521,
 * 3556-3576): from the running sums of ucat (usum), of uv|vw|wu (u1sum) and
 * of uu|vv|ww (u2sum) over N = ti - 1 steps (3P device fields each), per
 * element (n_elmt device doubles each) the mean over its nodes of
 * ustar_avg^2 (the wall model of opt, or noslip, on the mean velocity) and of
 * the three contravariant Reynolds stresses of csi-eta, eta-zet and csi-zet
 * (Contravariant_Reynolds_stress, ibm.c:3285-3292). */
int vfs_ib_averages(vfs_ctx *ctx, const vfs_ib_surface *surf, const vfs_ib_node *nodes, int n,
                    const vfs_ib_options *opt, const double *usum, const double *u1sum, const double *u2sum, double N,
                    double *mean_shear, double *rs1, double *rs2, double *rs3);

/* The flux forcing of k-periodic channels without dpdz_set (momentum.c:
 * 387-398): Pressure_Gradient adds dz (mean_k_flux - inlet_flux) / dt /
 * mean_k_area to the k-face gradient.  vfs_set_inlet_flux sets -flux
 * (main.c:198, 2067; the reference's default -1).  vfs_calc_k_flux is
 * Calc_k_Flux (rhs2.c:1562-1645): the fluid k flux of each plane averaged
 * over the planes, with the mean fluid |zet| of the planes of
 * Calc_Inlet_Area (bcs.c:254-278); both stay on the device for
 * Pressure_Gradient, and are returned when the host pointers are not NULL
 * (then the stream is synchronized).  Each plane is summed in a fixed tree
 * order, the planes in k order. */
int vfs_set_inlet_flux(vfs_ctx *ctx, double inlet_flux);
int vfs_calc_k_flux(vfs_ctx *ctx, const double *ucont, double *mean_k_flux, double *mean_k_area);
/* Pressure_Gradient (momentum.c:197-440). p: in/out (periodic ghosts).  A
 * k-periodic context without dpdz_set needs vfs_calc_k_flux first. */
int vfs_pressure_gradient(vfs_ctx *ctx, double *p, double *dP);
/* IB_BC (momentum.c:1982-2195). ucont in/out, ucat in/out. */
int vfs_ib_bc(vfs_ctx *ctx, double *ucont, double *ucat);
/* Formfunction_2 (momentum.c:456-1979): rhs += scale * F(ucont, ucat, nu_t)
 * on interior faces, masked faces and boundary planes set to zero.  ucont is
 * in/out (its periodic ghosts are rewritten as in momentum.c:642-669).
 * nu_t may be NULL when les == 0. */
int vfs_formfunction2(vfs_ctx *ctx, double *ucont, const double *ucat,
                      const double *nu_t, double *rhs, double scale);
/* FormFunction_SNES (momentum.c:2197-2301): rhs = -(U - U_o)/dt
 * + 0.5 FF2(U) + 0.5 rhs_o - dP.  U is not modified; lucont receives the
 * ghosted, IB-corrected flux the RHS was evaluated on; ucat in/out. */
int vfs_formfunction_snes(vfs_ctx *ctx, const double *U, const double *U_o,
                          const double *rhs_o, const double *dP, const double *nu_t,
                          double *lucont, double *ucat, double *rhs);
/* One stage of RungeKutta (rhs.c:1637-1659):
 * Contra2Cart; U = alfa*dt*st*(FF2(U) - dP) + U0; Contra2Cart.
 * rhs_scratch receives FF2(U) - dP. */
int vfs_rk_stage(vfs_ctx *ctx, const double *u0, double *ucont, double *ucat,
                 const double *nu_t, const double *dP, double alfa, double *rhs_scratch);
/* Boundary work of RungeKutta around its stages (rhs.c:1604-1614,
 * 1685-1686).  Pointers the configuration does not need may be NULL. */
typedef struct {
    const vfs_inflow_options *inflow;   /* inlet at k = 0 (bctype[4] = 5): InflowFlux */
    const double *coor;                 /* node coordinates (the inflow profile's y) */
    const double *uin_plane;            /* inletprofile 0: mx*my speeds, else NULL */
    double *ubcs;                       /* 3P boundary velocity, in/out, with an open k end */
    const vfs_ib_surface *surf;         /* immersed body: ibm_interpolation_advanced first */
    const vfs_ib_node *nodes;           /* its node list (vfs_ib_search) */
    int n_nodes;
    int add_body_force;                 /* 1: each stage also adds the context's body force
                                           (vfs_set_body_force); the reference adds F_eul only
                                           in FormFunction_SNES (momentum.c:2288), so 0 is the
                                           reference's explicit step */
    const vfs_ib_options *ib;           /* NULL: the defaults of vfs_ib_interpolate */
    double *p, *ustar;                  /* P fields the reconstruction updates */
} vfs_rk_bc;
/* RungeKutta (rhs.c:1578-1700): InflowFlux, FormBCS's outlet branch with
 * outflow_scale = 0 and the IB reconstruction, then four stages
 * (alfa = 1/4, 1/3, 1/2, 1; U0 = the flux after the boundary work, returned
 * in u0), then FormBCS's outlet branch with outflow_scale = 1.  ucont, ucat
 * in/out; u0, rhs: 3P scratch (rhs ends holding the last stage's FF2 - dP);
 * *flux_in / *flux_out (host, may be NULL): the inflow and the outflow
 * before the final correction.  bc may be NULL for a periodic, body-free
 * context.  Synchronizes the stream when a boundary routine runs. */
int vfs_runge_kutta(vfs_ctx *ctx, const vfs_rk_bc *bc, double *ucont, double *ucat, const double *nu_t,
                    const double *dP, double *u0, double *rhs, double *flux_in, double *flux_out);
/* Calc_Minimum_dt (momentum.c:104-195): CFL-1 time step. */
int vfs_min_dt(vfs_ctx *ctx, const double *ucont, double *dt_out);

/* ---- derivatives of the implicit residual ----------------------------------
 * F(U) is the FormFunction_SNES residual (momentum.c:2197-2301) with U_o,
 * rhs_o, dP and nu_t held fixed, and the ucat of cells with nvert >= 0.1
 * taken from `ucat` (frozen, as inside the reference's SNES solve).
 * J = dF/dU replaces the finite-difference MatMFFD product of
 * Implicit_MatrixFree (implicitsolver.c:4222-4236).  No argument is modified
 * except the output; v and jtv must not alias.  Contexts with a
 * wall-function wall (bctype -1) are rejected: the wall model is not
 * differentiated. */
/* jw = J w */
int vfs_rhs_jvp(vfs_ctx *ctx, const double *U, const double *ucat, const double *nu_t,
                const double *w, double *jw);
/* jtv = J^T v */
int vfs_rhs_vjp(vfs_ctx *ctx, const double *U, const double *ucat, const double *nu_t,
                const double *v, double *jtv);
/* Implicit_MatrixFree (implicitsolver.c:4196-4441): solves F(U) = 0 by
 * Newton's method to ||F|| <= rtol ||F(U_in)|| (rtol = imp_free_tol, 1e-4 by
 * default in main.c:212; at most max_newton steps, 50 in the reference),
 * each step an unpreconditioned restarted GMRES(restart) on J dU = -F with
 * the exact JVP and Eisenstat-Walker forcing terms (at most max_krylov
 * iterations, 1000 in the reference), then a backtracking line search on
 * ||F||.  U (in/out) returns with zero no-slip boundary fluxes; ucat (in/out)
 * returns Contra2Cart of the solution with the frozen cells unchanged.
 * Reports the Newton and total Krylov iteration counts and the final ||F||. */
int vfs_implicit_matrixfree(vfs_ctx *ctx, double *U, const double *U_o, const double *rhs_o,
                            const double *dP, const double *nu_t, double *ucat, double rtol,
                            int max_newton, int restart, int max_krylov, int *newton_its,
                            int *krylov_its, double *fnorm);

/* ---- pressure projection ------------------------------------------------ */
/* setup_lidx2 + PoissonLHSNew (poisson.c:182-329, 971-1682): numbers the
 * unknowns (returned in *n_unknowns) and builds the 19-point coefficients
 * (with the 1e-10 drop rule) and the multigrid hierarchy, from the context's
 * metrics and nvert.  Returns -8 when no cell is fluid (the system is
 * empty). */
int vfs_poisson_setup(vfs_ctx *ctx, int *n_unknowns);
/* Unknown numbering on the DA layout (gid, -1 where no unknown). */
int vfs_poisson_gid(vfs_ctx *ctx, double *gid);
/* The 19 coefficients per cell after the drop rule, slot-major
 * coef[m*P + cell], slot order of poisson.c:951-969. */
int vfs_poisson_coefficients(vfs_ctx *ctx, double *coef);
/* y = A x on the DA layout (FP64 operator). */
int vfs_poisson_apply(vfs_ctx *ctx, const double *x, double *y);
/* VolumeFlux (poisson.c:3570-3910 when immersed, 3306-3500 otherwise):
 * net IB flux correction of ucont in place; flux/area after correction. */
int vfs_volume_flux(vfs_ctx *ctx, double *ucont, double *flux, double *area);
/* PoissonRHS2_hypre + Remove_Nullspace_Scale (poisson_hypre.c:257-399,
 * 110-192): b on the DA layout, zero outside the unknowns. */
int vfs_poisson_rhs(vfs_ctx *ctx, const double *ucont, double *b);
/* Krylov + multigrid solve of A phi = b to ||b - A phi||/||b|| <= rtol in the
 * FP64 operator, then Remove_Nullspace and Convert_Phi2_Phi
 * (poisson_hypre.c:498-630, 194-255; poisson.c:92-180).
 * precision 0: FP64 throughout; 1: FP32 stencil/smoother/Krylov inside an
 * FP64 defect-correction loop (FP64 residual, dot products, updates).
 * phi: in = initial guess on the DA layout, out = solution (ghosts set).
 * Returns 0 when converged, 1 when maxit was reached above rtol (phi, iters
 * and relres hold the last iterate), -6 when b or the residual is not finite
 * (NaN / Inf).  The reference's HYPRE solve returns nothing in either case. */
int vfs_poisson_solve(vfs_ctx *ctx, const double *b, double *phi, double rtol,
                      int precision, int maxit, int *iters, double *relres);
/* UpdatePressure (poisson.c:3144-3304): p, phi in/out. */
int vfs_update_pressure(vfs_ctx *ctx, double *p, double *phi);
/* Projection (poisson.c:2707-3073) without the trailing Contra2Cart. */
int vfs_projection(vfs_ctx *ctx, const double *phi, double *ucont);
/* Divergence (main.c:1047-1177): max |div U| * aj over unmasked cells;
 * div (may be NULL) receives the per-cell field. */
int vfs_divergence(vfs_ctx *ctx, const double *ucont, double *div, double *maxdiv);

/* ---- reference file formats (host I/O) ------------------------------------
 * Replaces the grid reading of MG_Initial (init.c:316-436): format 0 text
 * grid.dat, 1 binary grid.dat (-binary 1), 2 tensor-product xyz.dat (-xyz 1).
 * Single-block files only.  Node coordinates value / cl * L_dim
 * (cl = -chact_leng, init.c:289-291) are written to coor in the layout
 * vfs_build_metrics takes: (IM+1)*(JM+1)*(KM+1) Cmpnts, node (k,j,i) at
 * ((k*(JM+1) + j)*(IM+1) + i); the other points are left untouched.
 * coor == NULL returns only dims = {IM, JM, KM}.  Host memory. */
int vfs_read_grid(const char *path, int format, double cl, double L_dim, int dims[3], double *coor);
/* PETSc binary vector files (big-endian class id 1211214, int32 length,
 * float64 values): what VecView / VecLoadIntoVector write and read for the
 * restart files (main.c:394-420, 799-847).  Host memory; reading requires the
 * file's length to equal n. */
int vfs_vec_write_petsc(const char *path, const double *v, long long n);
int vfs_vec_read_petsc(const char *path, double *v, long long n);
/* The restart set "<dir>/{vfield,ufield,pfield,nvfield}%06d_%1d.dat" of one
 * block at step ti: Ucont_P_Binary_Output (main.c:799-847, without the
 * optional averaging and Q fields) and Ucont_P_Binary_Input
 * (main.c:394-420, which loads nvfield into Nvert_o).  Device arrays. */
int vfs_restart_write(vfs_ctx *ctx, const char *dir, int ti, int block, const double *ucont,
                      const double *ucat, const double *p, const double *nvert);
int vfs_restart_read(vfs_ctx *ctx, const char *dir, int ti, int block, double *ucont, double *ucat,
                     double *p, double *nvert_o);
int vfs_restart_path(char *out, size_t cap, const char *dir, const char *stem, int ti, int block);

/* ---- synthetic workloads (host-side generators) -------------------------- */
double vfs_wl_uniform(uint64_t seed, uint64_t counter);
int vfs_wl_channel_grid(int IM, int JM, int KM, double Lx, double Ly, double Lz,
                        double beta, double amp, double *coor_host);
int vfs_wl_loglaw_k(int mx, int my, int mz, const int per[3], double Ly, double y0, double noise, uint64_t seed,
                    const double *cent, const double *nvert, const double *icsi, const double *jeta,
                    const double *kzet, double *ucat, double *ucont);
int vfs_wl_poiseuille(int mx, int my, int mz, const int per[3], double Ly, double noise,
                      uint64_t seed, const double *cent, const double *nvert,
                      const double *icsi, const double *jeta, const double *kzet,
                      double *ucat, double *ucont);
int vfs_wl_turbulent(int mx, int my, int mz, const int per[3], double Lx, double Ly, double Lz,
                     int nmodes, int nmax, double urms, uint64_t seed, const double *cent,
                     const double *nvert, const double *icsi, const double *jeta,
                     const double *kzet, double *ucat, double *ucont);
int vfs_wl_random_scalar(int mx, int my, int mz, double amp, uint64_t seed,
                         const double *nvert, const double *aj, double *out);
int vfs_wl_random_vector(int mx, int my, int mz, double amp, uint64_t seed, double *out);

#ifdef __cplusplus
}
#endif
#endif
