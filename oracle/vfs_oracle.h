/*
 * CPU oracle for the VFS-Geophysics fractional-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded FP64
 * restatement of the reference routines, written from reading the reference
 * source as text.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product path
 * (libvfsgpu.so) never links, loads or calls anything in this directory.
 *
 * Layout contract (identical to the reference DMDA single-rank arrays):
 *   sizes mx = IM+1, my = JM+1, mz = KM+1 (IM = node count),
 *   scalar field  s[(k*my + j)*mx + i]
 *   vector field  v[3*((k*my + j)*mx + i) + c],  c = x,y,z  (Cmpnts AoS)
 *   cells 1..m-2 are interior, 0 and m-1 are ghosts; periodic directions use
 *   the i_periodic convention (ghost 0 == cell m-2, ghost m-1 == cell 1).
 */
#ifndef VFS_ORACLE_H
#define VFS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int mx, my, mz;
    int i_periodic, j_periodic, k_periodic;
    int bctype[6];
    int immersed;          /* 0: no IB, 1: CURVIB masks (poisson.c:3570 VolumeFlux) */
    int les;               /* 0/1: constant-Cs Smagorinsky                          */
    int second_order;      /* momentum.c -second_order                              */
    int dpdz_set;          /* 1: k-periodic forcing uses mean_pressure_gradient     */
    int weno;              /* 1: WENO3 convective face values (momentum.c:782-786)  */
    int inviscid;          /* -inv: WENO3 convection, 2nd-order divergence, no viscous term */
    int skew;              /* -skew: skew-symmetric convection (momentum.c:798-811, 1596-1608) */
    int clark;             /* -clark: Clark model stress in the viscous fluxes (momentum.c:907-925)
                              and its Germano term in the dynamic procedure (les.c:648) */
    double ren, dt, st;
    double mean_pressure_gradient;
} vo_cfg;

typedef struct {
    double *csi, *eta, *zet, *aj;
    double *icsi, *ieta, *izet, *iaj;
    double *jcsi, *jeta, *jzet, *jaj;
    double *kcsi, *keta, *kzet, *kaj;
    double *cent;
} vo_geom;

/* threads of the parallel loops of vo_formfunction2 (OpenMP over k; the
 * results are the same for any count) */
int vo_threads(void);

/* metrics.c:4-1346 (NEWMETRIC). coor: node coordinates on the DA layout
 * (nodes i<mx-1, j<my-1, k<mz-1). All geom arrays must be zero-initialised. */
int vo_form_metrics(const vo_cfg *c, const double *coor, vo_geom *g);

/* ibm.c:319-341, 508-536: nvert mask from a sphere signed distance at cell
 * centres: 3 solid, 1 fluid cell with a solid cell in its 3x3x3 box, 0 fluid. */
int vo_nvert_sphere(const vo_cfg *c, const vo_geom *g, double cx, double cy,
                    double cz, double r, double *nvert);

/* rhs.c:57-741 Contra2Cart_2.  ucont: in/out (periodic ghosts), ucat: in/out. */
int vo_contra2cart(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   double *ucont, double *ucat);

/* The same, also writing the friction velocity of the wall-function cells
 * (bctype -1, rhs.c:303-401) into ustar (user->lUstar); other points untouched. */
int vo_contra2cart_ustar(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         double *ucont, double *ucat, double *ustar);

/* wallfunction.c: the Cabot wall model of bctype -1.  The y+ integral table
 * of pre_integrate (291-323), integrate_F (325-371), the Newton solve of
 * find_utau_Cabot (389-404) and wall_function (18-43) with Ua = 0. */
const double *vo_wf_table(void);
double vo_wf_integrate_F(double nu, double utau, double yb);
double vo_wf_find_utau(double nu, double u, double y, double guess);
void vo_wall_function(double nu, double sc, double sb, const double Uc[3], double Ub[3], double *ustar,
                      double nx, double ny, double nz);

/* momentum.c:2014-2039: without an immersed body, the first cells of the
 * wall-function walls become IB nodes (nvert = 1). */
int vo_wall_mask(const vo_cfg *c, double *nvert);

/* les.c:1135-1252 Compute_eddy_viscosity_LES with a constant Cs (les.c:67-77). */
int vo_eddy_viscosity(const vo_cfg *c, const vo_geom *g, const double *nvert,
                      const double *ucat, double cs, double *nu_t);

/* The same with a per-cell Cs field (the dynamic procedure's lCs). */
int vo_eddy_viscosity_cs(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         const double *ucat, const double *cs, double *nu_t);

/* The options of Compute_Smagorinsky_Constant_1 (main.c:202-255, 1721-1772,
 * 2066): -les, -i/j/k_homo_filter, -testfilter_ik, -max_cs, and the step
 * counters of the fresh-start and -les 3 rules. */
typedef struct {
    int les;                 /* 1 constant, 2 dynamic, 3 dynamic with relaxation */
    int i_homo_filter, j_homo_filter, k_homo_filter;
    int testfilter_ik;
    double max_cs;
    int ti, tistart, rstart;
} vo_les_opt;

/* les.c:67-1133 Compute_Smagorinsky_Constant_1, with the Clark term when
 * c->clark is set (o NULL: -les 2 with the default options at ti = 2).  Cs = 0 while ti < 2
 * on a fresh start.  Cs: in (the previous Cs, read by -les 3) and out. */
int vo_smagorinsky_constant(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucat,
                            const vo_les_opt *o, double *Cs);

/* Its stages, exposed for their own tests.  rhs2.c:412-475 integrate_testfilter
 * and rhs2.c:388-410 integrate_testfilter_ik (val, w indexed [k][j][i] offset
 * + 1); pass 1 (les.c:188-304): grad (9 doubles a point: du/dx, du/dy, du/dz,
 * dv/dx, ...), |S| and the filtered velocity; pass 2 (306-706): LM and MM;
 * pass 3 with the homogeneous averages and the clipping (708-1047); the
 * final Cs smoothing (1049-1107, in place). */
double vo_testfilter(const double val[3][3][3], const double w[3][3][3]);
double vo_testfilter_ik(const double val[3][3][3]);
int vo_dsm_pass1(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucat, const vo_les_opt *o,
                 double *grad, double *sabs, double *ucat_f);
int vo_dsm_pass2(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucat, const vo_les_opt *o,
                 const double *grad, const double *sabs, const double *ucat_f, double *LM, double *MM);
int vo_dsm_pass3(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *LM, const double *MM,
                 const vo_les_opt *o, double *Cs);
int vo_dsm_smooth(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_les_opt *o, double *Cs);

/* momentum.c:197-440 Pressure_Gradient. p: in/out (periodic ghosts). */
int vo_pressure_gradient(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         double *p, double *dp);

/* the same with the flux forcing of a k-periodic context without
 * dpdz_set (momentum.c:387-398): dz (mean_k_flux - inlet_flux) / dt /
 * mean_k_area added to the k-face gradient */
int vo_pressure_gradient_kf(const vo_cfg *c, const vo_geom *g, const double *nvert, double *p, double *dp,
                            double mean_k_flux, double inlet_flux, double mean_k_area);
/* rhs2.c:1562-1645 Calc_k_Flux and the mean plane area of bcs.c:254-278 */
int vo_calc_k_flux(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucont, double *mean_k_flux,
                   double *mean_k_area);

/* momentum.c:1982-2195 IB_BC. ucont in/out, ucat in/out (periodic ghosts). */
int vo_ib_bc(const vo_cfg *c, const vo_geom *g, const double *nvert,
             double *ucont, double *ucat);

/* momentum.c:456-1979 Formfunction_2: rhs += scale * F(ucont, ucat).
 * ucont: in/out (periodic ghosts are rewritten, momentum.c:642-669). */
/* level.c:1734-1771 weno3 (exposed for its known-answer tests) */
double vo_weno3(double f0, double f1, double f2, double f3, double wavespeed);

int vo_formfunction2(const vo_cfg *c, const vo_geom *g, const double *nvert,
                     double *ucont, const double *ucat, const double *nu_t,
                     double *rhs, double scale);

/* momentum.c:2197-2301 FormFunction_SNES.  U: candidate Ucont (not modified);
 * lucont: scratch out (the ghosted/IB-corrected flux the RHS saw);
 * ucat: in/out (IB-node values are inputs). */
int vo_formfunction_snes(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         const double *U, const double *U_o, const double *rhs_o,
                         const double *dP, const double *nu_t, double *lucont,
                         double *ucat, double *rhs);

/* The same with the rotor/nacelle body force added last (momentum.c:2288);
 * f_eul NULL: none. */
int vo_formfunction_snes_f(const vo_cfg *c, const vo_geom *g, const double *nvert,
                           const double *U, const double *U_o, const double *rhs_o,
                           const double *dP, const double *nu_t, const double *f_eul, double *lucont,
                           double *ucat, double *rhs);

/* Actuator models.  rotor_model.c:6042-6200 discrete delta functions
 * (df 0: dfunc_2h, 6: dfunc_nh(n), 7: dfunc_exp(n), 8: dfunc_s4h, other:
 * dfunc_sc4h); rotor_model.c:2590-2737 element search (closest node, widths,
 * index boxes); rotor_model.c:4792-5137 Calc_F_eul.  Element arrays are
 * blocks of n: cent x|y|z, F_lagr x|y|z, box i_min|i_max|j_min|j_max|k_min|k_max,
 * dh x|y|z. */
double vo_dfunc(int df, double r, double n);
int vo_rotor_elements(const vo_cfg *c, const vo_geom *g, const double *coor, int n, const double *cent, int itpwidth,
                      int *box, double *dh);
int vo_calc_f_eul(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *coor, int n,
                  const double *cent, const double *F, const double *dA, const int *box, int df, double halfwidth,
                  int width_mode, const double *dh_fixed, double *f_eul);

/* momentum.c:104-195 Calc_Minimum_dt: CFL-1 time step. */
double vo_min_dt(const vo_cfg *c, const vo_geom *g, const double *ucont);

/* ---- pressure projection ---- */
/* poisson.c:182-329 setup_lidx2; returns the number of unknowns. gid: DA layout. */
int vo_setup_gid(const vo_cfg *c, const double *nvert, double *gid);

/* poisson.c:971-1682 PoissonLHSNew as a CSR matrix over the unknowns.
 * rowptr[n+1], col[19n], val[19n] preallocated; returns nnz. */
int vo_poisson_lhs(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   const double *gid, int *rowptr, int *col, double *val);

/* The 19 stencil coefficients per cell before the drop rule (index map
 * poisson.c:951-969).  coef[19*P] on the DA layout, zero where no row. */
int vo_poisson_stencil(const vo_cfg *c, const vo_geom *g, const double *nvert,
                       double *coef);

/* poisson_hypre.c:257-399 PoissonRHS2_hypre + Remove_Nullspace_Scale (110-192)
 * b on the DA layout (zero outside unknowns). */
int vo_poisson_rhs(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   const double *ucont, double *b);

/* poisson.c:3570-3910 VolumeFlux (immersed==1, flg=0). ucont in/out. */
int vo_volume_flux(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   double *ucont, double *flux_out, double *area_out);

/* FP64 restarted GMRES on the assembled CSR to rtol, then Remove_Nullspace
 * (poisson_hypre.c:194-255) and Convert_Phi2_Phi (poisson.c:92-180).
 * b, phi on the DA layout.  Returns iterations, *relres the true residual. */
int vo_poisson_solve(const vo_cfg *c, const double *nvert, const double *gid,
                     const int *rowptr, const int *col, const double *val,
                     const double *b, double rtol, int maxit, double *phi,
                     double *relres);

/* y = A x on the DA layout with the assembled CSR. */
int vo_csr_apply(const vo_cfg *c, const double *gid, const int *rowptr,
                 const int *col, const double *val, const double *x, double *y);

/* poisson.c:3144-3304 UpdatePressure.  p, phi in/out (periodic ghosts). */
int vo_update_pressure(const vo_cfg *c, const double *nvert, double *p, double *phi);

/* poisson.c:2707-3073 Projection (without the trailing Contra2Cart).
 * phi read-only (periodic ghosts must be set), ucont in/out. */
int vo_projection(const vo_cfg *c, const vo_geom *g, const double *nvert,
                  const double *phi, double *ucont);

/* main.c:1047-1177 Divergence: returns max |div U| * aj. */
double vo_divergence(const vo_cfg *c, const vo_geom *g, const double *nvert,
                     const double *ucont, double *div);

/* rhs.c:1637-1659 one Runge-Kutta stage (alfa = 0.25, 1/3, 0.5, 1 in turn):
 * U = alfa*dt*st*(FF2(U) - dP) + U0.  ucont, ucat in/out. */
int vo_rk_stage(const vo_cfg *c, const vo_geom *g, const double *nvert,
                const double *u0, double *ucont, double *ucat, const double *nu_t,
                const double *dP, double alfa, double *rhs_scratch);

/* A triangulated body (IBMNodes): vertices x|y|z_bp[n_v], elements
 * nv1..nv3[n_elmt] with unit normals nf pointing into the fluid, vertex
 * velocities u (3*n_v, Cmpnts; NULL = at rest). */
typedef struct {
    int n_v, n_elmt;
    const double *x_bp, *y_bp, *z_bp;
    const int *nv1, *nv2, *nv3;
    const double *nf_x, *nf_y, *nf_z;
    const double *u;
} vo_ib_surface;

/* One IB node and its reconstruction geometry (IBMInfo, variables.h:126-172). */
typedef struct {
    int i, j, k, cell, imode;
    int i1, j1, k1, i2, j2, k2, i3, j3, k3;
    double cr1, cr2, cr3, cs1, cs2, cs3, d_s, d_i;
    double pmin[3];
} vo_ib_node;

/* ibm.c:1276-1364 BoundingSphere of element e */
void vo_ib_bounding_sphere(const vo_ib_surface *s, int e, double q[3], double *rad);
/* cell of point n (0..8) of interception patch pl (0..5) around (i, j, k) */
void vo_ib_patch_point(int pl, int n, int i, int j, int k, int *ii, int *jj, int *kk);

/* ibm.c:366-497 ibm_search_advanced, the node list: nearestcell1 and
 * InterceptionPoint for every IB node, in the list's order (last cell found
 * first).  -3: a node with no element or no interception; -4: cap too small. */
int vo_ib_search(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_ib_surface *s, vo_ib_node *nodes,
                 int cap, int *n_nodes);

/* ibm.c:3294-3698 ibm_interpolation_advanced (no averaging, levelset, FSI or
 * IB wall model): ucat and p of the nodes, the element means of ustar^2 and
 * p, the fluxes of the IB faces, then Contra2Cart.  ucont, ucat, p, ustar
 * (a field) in/out; count, shear, pressure: n_elmt each. */
int vo_ib_interpolate(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_ib_surface *s,
                      const vo_ib_node *nodes, int n, double *ucont, double *ucat, double *p, double *ustar,
                      int *count, double *shear, double *pressure);

/* options of ibm_interpolation_advanced: moving (movefsi / rotatefsi: no
 * fluid-support guard), wallfunction (-wallfunction with ti > 0: the
 * rough-wall law of wall_function_roughness_a with roughness ks; 2: with
 * -roughness on body 0, wall_function_roughness_loglaw with roughness
 * "roughness" at the nodes, law 1 in the averages) */
typedef struct {
    int moving, wallfunction;
    double ks, roughness;
} vo_ib_opt;
int vo_ib_interpolate_opt(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_ib_surface *s,
                          const vo_ib_node *nodes, int n, const vo_ib_opt *opt, double *ucont, double *ucat, double *p,
                          double *ustar, int *count, double *shear, double *pressure);
/* ibm.c:3481-3521, 3556-3576: the averaging statistics per element */
int vo_ib_averages(const vo_cfg *c, const vo_geom *g, const vo_ib_surface *s, const vo_ib_node *nodes, int n,
                   const vo_ib_opt *opt, const double *usum, const double *u1sum, const double *u2sum, double N,
                   double *mean_shear, double *rs1, double *rs2, double *rs3);
/* sediment_transport.c:29-68: the rough-wall law and its Newton inverse */
double vo_hydset_u(double nu, double y, double utau, double ks);
double vo_hydset_utau(double nu, double u, double y, double guess, double ks);
/* wallfunction.c:221-231 u_loglaw_roughness and 423-437
 * find_utau_loglaw_roughness (wallfunction == 2 with roughness ks) */
double vo_roughlog_u(double y, double utau, double ks);
double vo_roughlog_utau(double u, double y, double guess, double ks);

/* ibm_search_advanced's marking (ibm.c:110-360, 508-516): nvert of one
 * triangulated body by ray casting (point_cell_advanced, ibm.c:1068-1183),
 * the 3x3x3 dilation and the relabelling to 3 / 1.  reset: the first body
 * of a search (nvert < 5 cleared first).  -3: a ray never came clear. */
int vo_nvert_body(const vo_cfg *c, const vo_geom *g, double *nvert, const vo_ib_surface *s, int reset);
/* the same with -thin (ibm.c:281-317): cells around elements the box
 * diagonals of the cell centres cut are marked too */
int vo_nvert_body_thin(const vo_cfg *c, const vo_geom *g, double *nvert, const vo_ib_surface *s, int reset, int thin);
/* the first rand() after srand(seed), scaled to [-0.5, 0.5) */
double vo_ib_ray_jitter(int seed);

/* Open k boundaries (an inlet at k = 0, an outlet at k = mz-1).
 * inletprofile 1/2: uniform (1, or inlet_flux / inlet_area when inlet_flux
 * > 0); 10: the power law of bcs.c:432-437; 21: 1.5 y (2 - y); anything
 * else: the caller's uin_plane[j*mx + i]. */
typedef struct {
    int inletprofile;
    int reserved;
    double inlet_flux;
    double inlet_area;
} vo_inflow_opt;
/* rhs.c:34-55 */
void vo_contra2cart_single(const double *csi, const double *eta, const double *zet, const double *uc, double *u);
/* bcs.c:216-328 Calc_Inlet_Area: k_area[mz], its mean over 1..mz-2, and the
 * fluid area of the k = 0 plane */
int vo_inlet_area(const vo_cfg *c, const vo_geom *g, const double *nvert, double *k_area, double *mean_k_area,
                  double *inlet_area);
/* bcs.c:329-706 InflowFlux (bctype[4] = 5, no levelset, no fluctuations) */
int vo_inflow_flux(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *coor, const vo_inflow_opt *o,
                   const double *uin_plane, double *ucont, double *ucat, double *ubcs, double *flux_in);
/* bcs.c:801-870 OutflowFlux (bctype[5] = 4) */
int vo_outflow_flux(const vo_cfg *c, const double *nvert, const double *ucont, double *flux_out);
/* bcs.c:2285-2418 FormBCS, outlet (bctype[5] = 4, no levelset); scale is
 * outflow_scale */
int vo_outlet_correction(const vo_cfg *c, const vo_geom *g, const double *nvert, double flux_in, int scale,
                         const double *ucat, double *ucont, double *ubcs, double *flux_out, double *area);

/* rhs.c:1578-1700 RungeKutta with its boundary work: InflowFlux (inlet at
 * k = 0), FormBCS's outlet branch with outflow_scale 0, the IB
 * reconstruction (when s is given), the four stages from U0 = the flux after
 * that work (returned in u0), and FormBCS's outlet branch with
 * outflow_scale 1.  f_eul (NULL: none) is added in every stage.  ubcs: 3P,
 * needed with an open k end; flux_in / flux_out may be NULL. */
int vo_runge_kutta(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *coor,
                   const vo_inflow_opt *inflow, const double *uin_plane, double *ubcs, const vo_ib_surface *s,
                   const vo_ib_node *nodes, int n_nodes, const vo_ib_opt *ib, double *p, double *ustar,
                   const double *f_eul, double *ucont, double *ucat, const double *nu_t, const double *dP, double *u0,
                   double *rhs, double *flux_in, double *flux_out);

#ifdef __cplusplus
}
#endif
#endif
