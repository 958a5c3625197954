"""One fractional step of Flow_Solver's explicit path (solvers.c:590-1390)
through the C-ABI, in the reference's order:

  1. on the first step, the IB reconstruction and IB_BC (solvers.c:599-610);
  2. Calc_k_Flux for the flux forcing of a k-periodic channel without
     dpdz_set (solvers.c:644); LES: Contra2Cart and the eddy viscosity
     (solvers.c:647-654);
  3. Pressure_Gradient (solvers.c:786);
  4. RungeKutta with its boundary work: InflowFlux, FormBCS (outflow_scale
     0), the IB reconstruction, four stages, FormBCS (outflow_scale 1)
     (rhs.c:1578-1700, vfs_runge_kutta);
  5. PoissonSolver: VolumeFlux, PoissonRHS2 with the null-space scaling, the
     solve, Remove_Nullspace (poisson_hypre.c:498-630);
  6. UpdatePressure and Projection (solvers.c:1336-1345);
  7. the IB reconstruction, Divergence, IB_BC and Contra2Cart
     (solvers.c:1362-1385).
"""
from . import InflowOptions, IbOptions

__all__ = ["FlowSolver"]


class FlowSolver:
    """The state of one block (the reference's UserCtx fields the step
    touches) on the device, and the step.

    ctx: a Ctx with metrics and mask set.  ucont, ucat: DeviceArrays (3P),
    owned by the solver after construction.  coor: the node coordinates
    (DeviceArray) when the context has an inlet.  inflow: InflowOptions
    (default: uniform inflow); uin_plane: a DeviceArray of mx*my speeds for
    inletprofile 0.  body: (IbSurface, nodes DeviceArray, n) for the CURVIB
    reconstruction of an immersed context.  cs: the Smagorinsky constant of
    -les 1.  precision: 0 FP64 Poisson, 1 FP32 inside FP64 refinement.
    inlet_flux: -flux, the target k flux of the forcing of a k-periodic
    context without dpdz_set."""

    def __init__(self, ctx, ucont, ucat, coor=None, inflow=None, uin_plane=None, body=None, ib=None, cs=0.01,
                 precision=0, rtol=1e-8, maxit=1000, add_body_force=False, inlet_flux=-1.0):
        self.ctx, self.cfg = ctx, ctx.cfgd
        self.uc, self.u = ucont, ucat
        self.coor, self.uin_plane = coor, uin_plane
        self.open_k = not self.cfg["k_periodic"] and (self.cfg["bctype"][4] == 5 or self.cfg["bctype"][5] == 4)
        if self.cfg["bctype"][4] == 5 and inflow is None:
            inflow = InflowOptions(1, 0, -1.0, 0.0)
        self.inflow = inflow
        self.body, self.ib = body, ib
        self.cs, self.precision, self.rtol, self.maxit = cs, precision, rtol, maxit
        self.add_body_force = add_body_force
        self.p, self.nut, self.ustar = ctx.sca(), ctx.sca(), ctx.sca()
        self.dp, self.u0, self.rhs = ctx.vec(), ctx.vec(), ctx.vec()
        self.ubcs = ctx.vec() if self.open_k else None
        self.phi, self.b = ctx.sca(), ctx.sca()
        self.flux_forcing = bool(self.cfg["k_periodic"] and not self.cfg["dpdz_set"])
        if self.flux_forcing:
            ctx.set_inlet_flux(inlet_flux)
        self.ti = 0

    def _reconstruct(self):
        if self.cfg["immersed"] and self.body is not None:
            surf, nodes, n = self.body
            o = self.ib or IbOptions(0, 0, 0.0, 0.0)
            self.ctx.ib_interpolate(surf, nodes, n, self.uc, self.u, self.p, self.ustar, stats=False,
                                    moving=bool(o.moving), wallfunction=bool(o.wallfunction), ks=o.ks,
                                    roughness=o.roughness if o.wallfunction == 2 else None)

    def step(self):
        ctx = self.ctx
        out = {}
        if self.ti == 0:
            self._reconstruct()
            ctx.ib_bc(self.uc, self.u)
        if self.flux_forcing:
            out["mean_k_flux"], _ = ctx.calc_k_flux(self.uc)
        if self.cfg["les"]:
            ctx.contra2cart(self.uc, self.u)
            ctx.eddy_viscosity(self.u, self.cs, self.nut)
        ctx.pressure_gradient(self.p, self.dp)
        sb = self.body if (self.cfg["immersed"] and self.body is not None) else (None, None, 0)
        out["flux_in"], out["flux_out"] = ctx.runge_kutta(
            self.uc, self.u, self.nut if self.cfg["les"] else None, self.dp, self.u0, self.rhs, inflow=self.inflow,
            coor=self.coor, uin_plane=self.uin_plane, ubcs=self.ubcs, surface=sb[0], nodes=sb[1], n_nodes=sb[2],
            ib=self.ib, p=self.p, ustar=self.ustar, add_body_force=self.add_body_force)
        if self.open_k:
            out["flux_out_corrected"] = ctx.outflow_flux(self.uc)
        ctx.volume_flux(self.uc)
        ctx.poisson_rhs(self.uc, self.b)
        out["iters"], out["relres"] = ctx.poisson_solve(self.b, self.phi, rtol=self.rtol, precision=self.precision,
                                                        maxit=self.maxit)
        ctx.update_pressure(self.p, self.phi)
        ctx.projection(self.phi, self.uc)
        self._reconstruct()
        out["maxdiv"] = ctx.divergence(self.uc)
        ctx.ib_bc(self.uc, self.u)
        ctx.contra2cart(self.uc, self.u)
        self.ti += 1
        return out
