#!/usr/bin/env python3
"""Benchmark of the fractional-step hot path on one MI355X.

Workload (default): the 256x128x128 stretched curvilinear LES channel with an
immersed sphere (tanh beta=2, sinusoidal bottom, Re=5000, Cs^2=0.01), seeded
solenoidal k^-5/3 velocity.  One step = one fractional step in the
reference's order (solvers.c:648-1385): eddy viscosity, Pressure_Gradient,
4-stage RungeKutta momentum (4 Formfunction_2 evaluations), VolumeFlux,
Poisson RHS, Krylov-multigrid solve to ||b-A phi||/||b|| <= 1e-8 from a zero
initial guess, UpdatePressure, Projection, Contra2Cart, Divergence.

Prints one JSON line.  value = RHS throughput in Gcell/s of Formfunction_2
(FP64, interior cells / mean kernel time from HIP events on the library's
stream); the Poisson time-to-1e-8 in FP64 and in mixed FP32/FP64, the step
times of both modes, the roofline of the dominant kernel and the implicit
momentum solve at CFL 2 (Newton-Krylov with the exact JVP, plus the JVP and
VJP times), the dynamic Smagorinsky procedure (-les 2) with its eddy
viscosity and the actuator-disk force spreading are reported in the same
line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vfs-geophysics-3.1_amd"))

HBM_PEAK_GBS = 8000.0
RHS_BYTES_PER_CELL = {1: 424, 0: 416}   # Formfunction_2 compulsory traffic, LES on/off


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on the null stream (the library's default stream)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e):
        assert self.h.hipEventRecord(e, C.c_void_p(0)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def setup(name, N=None, poisson=True):
    import vfsgpu as V
    from vfsgpu import workloads as W
    cfg, spec = W.make_config(name, N)
    ctx = V.Ctx(cfg)
    coor = W.grid(spec)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    if spec["immersed"]:
        ctx.nvert_sphere(*W.SPHERE)
    nvert = ctx.nvert().host()
    metrics = {n: ctx.metric(n).host() for n in ("cent", "icsi", "jeta", "kzet")}
    ucat, ucont = W.fields(cfg, spec, metrics, nvert)
    n_unknowns = ctx.poisson_setup() if poisson else 0
    return cfg, spec, ctx, ucat, ucont, n_unknowns


def time_poisson_setup(ctx, reps):
    """Median wall time of vfs_poisson_setup: the unknown numbering, the
    19-point coefficients and the multigrid hierarchy, built again from the
    current mask (what a step pays when a moving body changes the mask)."""
    ts = []
    for _ in range(reps):
        ctx.sync()
        t0 = time.perf_counter()
        ctx.poisson_setup()
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    return round(float(np.median(ts)) * 1e3, 2)


class State:
    def __init__(self, ctx, ucat, ucont):
        import vfsgpu as V
        self.uc = V.DeviceArray.from_host(ucont)
        self.u = V.DeviceArray.from_host(ucat)
        self.u0 = ctx.vec()
        self.nut = ctx.sca()
        self.p = ctx.sca()
        self.dp = ctx.vec()
        self.rhs = ctx.vec()
        self.phi = ctx.sca()
        self.zero = ctx.sca()
        self.b = ctx.sca()


def fractional_step(ctx, cfg, s, precision, rtol=1e-8, timers=None):
    from vfsgpu import workloads as W
    if cfg["les"]:
        ctx.eddy_viscosity(s.u, W.CS_LES, s.nut)
    ctx.pressure_gradient(s.p, s.dp)
    s.u0.copy_from(s.uc)
    for alfa in (0.25, 1 / 3.0, 0.5, 1.0):
        ctx.rk_stage(s.u0, s.uc, s.u, s.nut, s.dp, alfa, s.rhs)
    ctx.volume_flux(s.uc)
    ctx.poisson_rhs(s.uc, s.b)
    ctx.sync()
    t0 = time.perf_counter()
    s.phi.copy_from(s.zero)
    it, rel = ctx.poisson_solve(s.b, s.phi, rtol=rtol, precision=precision, maxit=1000)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.update_pressure(s.p, s.phi)
    ctx.projection(s.phi, s.uc)
    ctx.contra2cart(s.uc, s.u)
    maxdiv = ctx.divergence(s.uc)
    if timers is not None:
        timers.append(dict(poisson_s=t1 - t0, iters=it, relres=rel, maxdiv=maxdiv))
    return it, rel, maxdiv


def accuracy_leg(ctx, s):
    """The seeded field projected once with each solve (FP64 and mixed, both
    to ||b - A phi|| / ||b|| <= 1e-8 in the FP64 operator): the relative
    difference of the projected velocity fluxes and of phi, mixed against
    FP64, and max |div U| after each projection."""
    import vfsgpu as V
    u0 = s.uc.host()
    U, phis, div = [], [], []
    for prec in (0, 1):
        uc = V.DeviceArray.from_host(u0)
        b, phi = ctx.sca(), ctx.sca()
        ctx.volume_flux(uc)
        ctx.poisson_rhs(uc, b)
        ctx.poisson_solve(b, phi, rtol=1e-8, precision=prec, maxit=1000)
        ctx.projection(phi, uc)
        div.append(ctx.divergence(uc))
        U.append(uc.host())
        phis.append(phi.host())
    return {"velocity_rel_diff_l2": float(np.linalg.norm(U[1] - U[0]) / np.linalg.norm(U[0])),
            "velocity_rel_diff_max": float(np.abs(U[1] - U[0]).max() / np.abs(U[0]).max()),
            "phi_rel_diff_l2": float(np.linalg.norm(phis[1] - phis[0]) / np.linalg.norm(phis[0])),
            "max_div_fp64": div[0], "max_div_mixed": div[1],
            "max_div_before": float(ctx.divergence(V.DeviceArray.from_host(u0)))}


def time_rhs(ctx, cfg, s, reps):
    """Mean Formfunction_2 time (ms) from HIP events over `reps` calls."""
    ev = HipEvents()
    a, b = ev.event(), ev.event()
    ctx.contra2cart(s.uc, s.u)
    for _ in range(3):
        ctx.formfunction2(s.uc, s.u, s.nut, s.rhs, 1.0)
    ctx.sync()
    ev.record(a)
    for _ in range(reps):
        ctx.formfunction2(s.uc, s.u, s.nut, s.rhs, 1.0)
    ev.record(b)
    return ev.elapsed_ms(a, b) / reps


def implicit_leg(ctx, cfg, spec, s):
    """Implicit_MatrixFree (implicitsolver.c:4196-4441) at CFL 2 on the full
    grid: dt = 2 x Calc_Minimum_dt, Newton-Krylov with the exact JVP of the
    FormFunction_SNES residual, rtol 1e-4 (imp_free_tol); plus the time of one
    JVP and one VJP against one Formfunction_2."""
    import vfsgpu as V
    from vfsgpu import workloads as W
    Uo = V.DeviceArray.from_host(s.uc.host())
    ctx.contra2cart(Uo, s.u)
    if cfg["les"]:
        ctx.eddy_viscosity(s.u, W.CS_LES, s.nut)
    dt = 2.0 * ctx.min_dt(Uo)
    cfg2 = dict(cfg, dt=dt)
    c2 = V.Ctx(cfg2)
    c2.build_metrics(V.DeviceArray.from_host(W.grid(spec)))
    if spec["immersed"]:
        c2.nvert_sphere(*W.SPHERE)
    rhs_o = c2.vec()
    c2.formfunction2(V.DeviceArray.from_host(Uo.host()), s.u, s.nut, rhs_o, 1.0)
    cent = c2.metric("cent").host().reshape(-1, 3)
    p = V.DeviceArray.from_host(0.01 * np.sin(cent[:, 0]) * np.cos(np.pi * cent[:, 1] / 2))
    dP = c2.vec()
    c2.pressure_gradient(p, dP)
    U = V.DeviceArray.from_host(Uo.host())
    ucat = V.DeviceArray.from_host(s.u.host())
    lu, uu, r0 = c2.vec(), V.DeviceArray.from_host(s.u.host()), c2.vec()
    c2.formfunction_snes(U, Uo, rhs_o, dP, s.nut, lu, uu, r0)
    f0 = float(np.linalg.norm(r0.host()))
    ev = HipEvents()
    w, jw = V.DeviceArray.from_host(W.unit_vector(cfg, 14001)), c2.vec()
    for _ in range(2):
        c2.rhs_jvp(U, ucat, s.nut, w, jw)
        c2.rhs_vjp(U, ucat, s.nut, w, jw)
    c2.sync()
    a, b, e = ev.event(), ev.event(), ev.event()
    ev.record(a)
    for _ in range(5):
        c2.rhs_jvp(U, ucat, s.nut, w, jw)
    ev.record(b)
    for _ in range(5):
        c2.rhs_vjp(U, ucat, s.nut, w, jw)
    ev.record(e)
    jvp_ms, vjp_ms = ev.elapsed_ms(a, b) / 5, ev.elapsed_ms(b, e) / 5
    c2.sync()
    t0 = time.perf_counter()
    ni, ki, fn = c2.implicit_matrixfree(U, Uo, rhs_o, dP, s.nut, ucat, rtol=1e-4, max_newton=50, restart=30,
                                        max_krylov=1000)
    c2.sync()
    t1 = time.perf_counter()
    c2.close()
    return {"dt_cfl2": dt, "newton_its": ni, "krylov_its": ki, "rel_residual": fn / f0,
            "solve_ms": round((t1 - t0) * 1e3, 2), "jvp_ms": round(jvp_ms, 3), "vjp_ms": round(vjp_ms, 3)}


# the dynamic procedure's algorithmic bytes a point: ucat, the cell metrics,
# aj and nvert read once, Cs written once (every intermediate kept on chip)
DSM_BYTES_PER_CELL = 120
# what the five kernels of this build move a point (each array once): staging
# 168, gradient pass 216, Germano pass 232, Cs pass 40, smoothing 32
DSM_STAGED_BYTES_PER_CELL = 688


def dynamic_les_leg(ctx, cfg, s, reps, cpu_budget):
    """Compute_Smagorinsky_Constant_1 with -les 2 (les.c:67-1133) and the eddy
    viscosity from its Cs field on the full grid: HIP-event time per call,
    and the oracle on a 128x64x64 sample of the same generator."""
    import vfsgpu as V
    from vfsgpu import workloads as W
    cs = ctx.sca()
    for _ in range(2):
        ctx.smagorinsky_constant(s.u, cs)
        ctx.eddy_viscosity_cs(s.u, cs, s.nut)
    ctx.sync()
    ev = HipEvents()
    a, b, e = ev.event(), ev.event(), ev.event()
    ev.record(a)
    for _ in range(reps):
        ctx.smagorinsky_constant(s.u, cs)
    ev.record(b)
    for _ in range(reps):
        ctx.eddy_viscosity_cs(s.u, cs, s.nut)
    ev.record(e)
    dsm_ms, nut_ms = ev.elapsed_ms(a, b) / reps, ev.elapsed_ms(b, e) / reps
    Ni, Nj, Nk = cfg["mx"] - 2, cfg["my"] - 2, cfg["mz"] - 2
    P = cfg["mx"] * cfg["my"] * cfg["mz"]
    h = cs.host()
    out = {"procedure_ms": round(dsm_ms, 4), "eddy_viscosity_ms": round(nut_ms, 4),
           "gcells": round(Ni * Nj * Nk / (dsm_ms * 1e-3) / 1e9, 3),
           "bytes_per_point": DSM_BYTES_PER_CELL,
           "achieved_gbs": round(DSM_BYTES_PER_CELL * P / (dsm_ms * 1e-3) / 1e9, 1),
           "frac_hbm": round(DSM_BYTES_PER_CELL * P / (dsm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "staged_bytes_per_point": DSM_STAGED_BYTES_PER_CELL,
           "staged_gbs": round(DSM_STAGED_BYTES_PER_CELL * P / (dsm_ms * 1e-3) / 1e9, 1),
           "cs_max": round(float(h.max()), 4), "cs_nonzero_frac": round(float((h > 0).mean()), 3)}
    if cpu_budget > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_py import Oracle
        N = (128, 64, 64)
        c2, spec = W.make_config("C2", N)
        o = Oracle(c2, W.grid(spec))
        o.nvert_sphere(*W.SPHERE)
        ucat, ucont = W.fields(c2, spec, o.g, o.nvert)
        o.contra2cart(ucont, ucat)
        reps_c, t = 0, 0.0
        while t < cpu_budget and reps_c < 100:
            t0 = time.perf_counter()
            o.smagorinsky_constant(ucat)
            t += time.perf_counter() - t0
            reps_c += 1
        out["cpu_oracle_gcells"] = round(N[0] * N[1] * N[2] * reps_c / t / 1e9, 5)
        out["cpu_sample"] = f"{reps_c} calls of the oracle at 128x64x64 cells, one thread ({t:.1f} s)"
    return out


def actuator_leg(ctx, cfg, spec, reps, with_cpu):
    """The actuator-disk force on the full grid: a ring of 512 elements
    (radius 0.5 about (Lx/4, 1, Lz/2), normal to x, thrust -1 per unit area)
    spread with the Gaussian delta function (deltafunc 7, halfwidth 4, the
    reference's default width): element search and Calc_F_eul times."""
    import vfsgpu as V
    from vfsgpu import workloads as W
    n = 512
    th = 2 * np.pi * np.arange(n) / n
    cent = np.stack([np.full(n, W.LX / 4), 1.0 + 0.5 * np.cos(th), W.LZ / 2 + 0.5 * np.sin(th)])
    F = np.stack([-np.ones(n), np.zeros(n), np.zeros(n)])
    dA = np.full(n, 2 * np.pi * 0.5 / n * 0.1)
    coor = W.grid(spec)
    dcoor, dcent = V.DeviceArray.from_host(coor), V.DeviceArray.from_host(cent.reshape(-1))
    dF, ddA, f = V.DeviceArray.from_host(F.reshape(-1)), V.DeviceArray.from_host(dA), ctx.vec()
    box, dh = ctx.rotor_elements(dcoor, dcent)
    ctx.calc_f_eul(dcoor, dcent, dF, ddA, box, f, deltafunc=7, halfwidth=4.0)
    ctx.sync()
    ev = HipEvents()
    a, b = ev.event(), ev.event()
    ev.record(a)
    for _ in range(reps):
        ctx.rotor_elements(dcoor, dcent)
    ev.record(b)
    search_ms = ev.elapsed_ms(a, b) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.calc_f_eul(dcoor, dcent, dF, ddA, box, f, deltafunc=7, halfwidth=4.0)
    ctx.sync()
    spread_ms = (time.perf_counter() - t0) / reps * 1e3
    h = f.host()
    out = {"elements": n, "search_ms": round(search_ms, 4), "spread_ms": round(spread_ms, 4),
           "nonzero_points": int(np.count_nonzero(h.reshape(-1, 3).any(axis=1)))}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_py import Oracle
        o = Oracle(cfg, coor)
        o.nvert_sphere(*W.SPHERE)
        t0 = time.perf_counter()
        obox, _ = o.rotor_elements(coor, cent)
        o.calc_f_eul(coor, cent, F, dA, obox, df=7, halfwidth=4.0)
        out["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
    return out


def ib_leg(ctx, cfg, spec, s, reps, with_cpu):
    """CURVIB reconstruction at the immersed sphere as 5120 triangles: the
    ray-casting mask and the node search (once per body position) and the
    per-step ibm_interpolation_advanced pass (node values, element means,
    IB-face fluxes, Contra2Cart), on copies of the state."""
    import vfsgpu as V
    from vfsgpu import workloads as W
    x, e, n = W.icosphere(*W.SPHERE, level=4)
    surf = V.IbSurface(x, e, n)
    # the mask of the triangulated sphere (ray casting); the analytic mask is
    # restored at the end
    ctx.nvert_body(surf)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.nvert_body(surf)
    mark_ms = (time.perf_counter() - t0) / reps * 1e3
    nodes, nn = ctx.ib_search(surf)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        nodes, nn = ctx.ib_search(surf)
    search_ms = (time.perf_counter() - t0) / reps * 1e3
    uc, u, p, ustar = ctx.vec(), ctx.vec(), ctx.sca(), ctx.sca()
    uc.copy_from(s.uc)
    u.copy_from(s.u)
    p.copy_from(s.p)
    ctx.ib_interpolate(surf, nodes, nn, uc, u, p, ustar)
    ctx.sync()
    ev = HipEvents()
    a, b = ev.event(), ev.event()
    ev.record(a)
    for _ in range(reps):
        ctx.ib_interpolate(surf, nodes, nn, uc, u, p, ustar)
    ev.record(b)
    interp_ms = ev.elapsed_ms(a, b) / reps
    ctx.nvert_sphere(*W.SPHERE)
    out = {"ib_nodes": nn, "elements": len(e), "marking_ms": round(mark_ms, 4), "search_ms": round(search_ms, 4),
           "interpolate_ms": round(interp_ms, 4)}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_py import IbBody, Oracle
        o = Oracle(cfg, W.grid(spec))
        body = IbBody(x, e, n)
        t0 = time.perf_counter()
        o.nvert_body(body)
        out["cpu_oracle_marking_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        t0 = time.perf_counter()
        onodes = o.ib_search(body)
        out["cpu_oracle_search_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        ucont, ucat = s.uc.host(), s.u.host()
        t0 = time.perf_counter()
        o.ib_interpolate(body, onodes, ucont, ucat, s.p.host(), np.zeros(o.P))
        out["cpu_oracle_interpolate_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
    return out


def turbine_share_leg(steps, warmup):
    """One GPU's share of the turbine-wake case (768 x 384 x 384 in 8 k
    slabs: 768 x 384 x 48), with its boundary set: a log-law inflow at k = 0
    (InflowFlux with the caller's profile), the outlet at k = mz-1, a
    wall-function bottom and LES.  The step is Flow_Solver's explicit path
    in the reference's order (vfsgpu.step.FlowSolver: RungeKutta with
    InflowFlux and FormBCS before the stages and the outlet correction after
    them), with the mixed-precision Poisson solve."""
    import vfsgpu as V
    from vfsgpu import workloads as W
    from vfsgpu.step import FlowSolver
    cfg, spec, ctx, ucat, ucont, nunk = setup("C5s")
    Ni, Nj, Nk = spec["N"]
    coor = W.grid(spec)
    dcoor = V.DeviceArray.from_host(coor)
    plane = V.DeviceArray.from_host(W.loglaw_inflow_plane(cfg, coor))
    _, _, inlet_area = ctx.inlet_area()
    fs = FlowSolver(ctx, V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), coor=dcoor,
                    inflow=V.InflowOptions(0, 0, -1.0, inlet_area), uin_plane=plane, cs=W.CS_LES, precision=1,
                    rtol=1e-8)
    for _ in range(warmup):
        fs.step()
    ctx.sync()
    t0 = time.perf_counter()
    out = [fs.step() for _ in range(steps)]
    ctx.sync()
    step_ms = (time.perf_counter() - t0) / steps * 1e3
    s = State(ctx, fs.u.host(), fs.uc.host())
    ctx.contra2cart(s.uc, s.u)
    ctx.eddy_viscosity(s.u, W.CS_LES, s.nut)
    rhs_ms = time_rhs(ctx, cfg, s, 10)
    cells = Ni * Nj * Nk
    return {"workload": f"turbine-wake share {Ni}x{Nj}x{Nk}: log-law inlet, outlet, wall-function bottom, LES",
            "step_ms_mixed": round(step_ms, 3), "poisson_iters": int(np.median([o["iters"] for o in out])),
            "max_relres": max(o["relres"] for o in out),
            "max_div_after_projection": max(o["maxdiv"] for o in out), "rhs_ms": round(rhs_ms, 4),
            "rhs_gcells": round(cells / (rhs_ms * 1e-3) / 1e9, 3),
            "flux_balance_after_outlet_correction": max(abs(o["flux_out_corrected"] - o["flux_in"]) / abs(o["flux_in"])
                                                        for o in out)}


def turbine_c5_leg(steps, warmup, precision=1):
    """The turbine-wake case at full size (768 x 384 x 384, 113 M cells):
    log-law inflow with fluctuations at k = 0 (InflowFlux with the caller's
    profile), the outlet at k = mz-1, a wall-function bottom, LES, and the
    actuator ring spread once with the Gaussian delta function and carried in
    every Runge-Kutta stage; FlowSolver steps with the mixed-precision
    Poisson solve to 1e-8.  Reports the step time and, for every step, that
    the fields stay finite, the outflow equals the inflow after the outlet
    correction, and the divergence left by the projection."""
    import vfsgpu as V
    from vfsgpu import workloads as W
    from vfsgpu.step import FlowSolver
    t_setup = time.perf_counter()
    log("[bench] C5: building the grid, metrics, fields and the Poisson hierarchy")
    cfg, spec, ctx, ucat, ucont, nunk = setup("C5")
    log(f"[bench] C5: {nunk} unknowns after {time.perf_counter() - t_setup:.1f} s")
    Ni, Nj, Nk = spec["N"]
    coor = W.grid(spec)
    dcoor = V.DeviceArray.from_host(coor)
    plane = V.DeviceArray.from_host(W.loglaw_inflow_plane(cfg, coor))
    del coor
    _, _, inlet_area = ctx.inlet_area()
    ring = W.ACTUATOR_RING
    cent, F, dA = W.actuator_ring()
    dcent = V.DeviceArray.from_host(cent.reshape(-1))
    box, _ = ctx.rotor_elements(dcoor, dcent)
    f_eul = ctx.vec()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.calc_f_eul(dcoor, dcent, V.DeviceArray.from_host(F.reshape(-1)), V.DeviceArray.from_host(dA), box, f_eul,
                   deltafunc=ring["deltafunc"], halfwidth=ring["halfwidth"])
    ctx.sync()
    spread_ms = (time.perf_counter() - t0) * 1e3
    ctx.set_body_force(f_eul)
    log(f"[bench] C5: actuator force spread in {spread_ms:.1f} ms")
    fs = FlowSolver(ctx, V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), coor=dcoor,
                    inflow=V.InflowOptions(0, 0, -1.0, inlet_area), uin_plane=plane, cs=W.CS_LES,
                    precision=precision, rtol=1e-8, add_body_force=True)
    del ucat, ucont
    setup_s = time.perf_counter() - t_setup
    for _ in range(warmup):
        fs.step()
    ctx.sync()
    t0 = time.perf_counter()
    out = []
    for _ in range(steps):
        out.append(fs.step())
        log(f"[bench] C5 step {fs.ti}: {out[-1]}")
    ctx.sync()
    step_ms = (time.perf_counter() - t0) / max(1, steps) * 1e3
    finite = all(np.isfinite(o["maxdiv"]) and np.isfinite(o["relres"]) for o in out)
    res = {"workload": f"turbine wake {Ni}x{Nj}x{Nk}: log-law inlet, outlet, wall-function bottom, LES, "
                       f"actuator ring of {cent.shape[1]} elements", "cells": Ni * Nj * Nk, "steps": steps,
           "warmup": warmup, "precision": "mixed" if precision else "fp64", "setup_s": round(setup_s, 1),
           "f_eul_spread_ms": round(spread_ms, 2), "step_ms": round(step_ms, 1),
           "poisson_iters": [o["iters"] for o in out], "max_relres": max(o["relres"] for o in out),
           "max_div_after_projection": max(o["maxdiv"] for o in out),
           "flux_balance_after_outlet_correction": max(abs(o["flux_out_corrected"] - o["flux_in"]) / abs(o["flux_in"])
                                                       for o in out),
           "finite": bool(finite and np.isfinite(fs.uc.host()).all())}
    return res


def traffic_from_profile(name):
    """HBM bytes per Formfunction_2 call from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE of the Formfunction_2 kernels, each
    scaled by the same counter on a device copy of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", f"formfunction2_{name}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_call\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def rhs_only(ctx, cfg, s, reps, calib_reps):
    """Short run for the counter passes: Formfunction_2 `reps` times, then a
    device-to-device copy of 3P doubles `calib_reps` times (the calibration
    kernel of known traffic)."""
    from vfsgpu import workloads as W
    import vfsgpu as V
    ctx.contra2cart(s.uc, s.u)
    if cfg["les"]:
        ctx.eddy_viscosity(s.u, W.CS_LES, s.nut)
    for _ in range(reps):
        ctx.formfunction2(s.uc, s.u, s.nut, s.rhs, 1.0)
    ctx.sync()
    nbytes = 3 * ctx.P * 8
    for _ in range(calib_reps):
        V._ck(V.lib().vfs_memcpy_dtod(s.rhs.ptr, s.u.ptr, C.c_size_t(nbytes)), "vfs_memcpy_dtod")
    ctx.sync()
    log(f"[bench] rhs-only: {reps} Formfunction_2 calls, {calib_reps} copies of {nbytes} bytes")


def cpu_baseline(name, budget_s):
    """The CPU oracle's Formfunction_2 on a bounded sample of the same
    generator: its point loops run over k on every host thread OpenMP gets
    (the results are the same as one thread's), and one thread for
    comparison."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    from oracle_py import Oracle
    from vfsgpu import workloads as W
    N = (128, 64, 64) if name != "C1" else (64, 64, 64)
    cfg, spec = W.make_config(name, N)
    coor = W.grid(spec)
    o = Oracle(cfg, coor)
    if spec["immersed"]:
        o.nvert_sphere(*W.SPHERE)
    ucat, ucont = W.fields(cfg, spec, o.g, o.nvert)
    o.contra2cart(ucont, ucat)
    nut = o.sca()
    if cfg["les"]:
        o.eddy_viscosity(ucat, W.CS_LES, nut)
    rhs = o.vec()
    cells = N[0] * N[1] * N[2]

    def timed(budget):
        reps, t = 0, 0.0
        while t < budget and reps < 5000:
            t0 = time.perf_counter()
            o.formfunction2(ucont, ucat, nut, rhs, 1.0)
            t += time.perf_counter() - t0
            reps += 1
        return reps, t
    threads = int(oracle_py.lib().vo_threads())
    reps, t = timed(0.7 * budget_s)
    out = dict(value=round(cells * reps / t / 1e9, 5), unit="Gcell/s", cores=threads, kind="port",
               sample=f"{reps} Formfunction_2 evaluations of the {name} generator at {N[0]}x{N[1]}x{N[2]} cells, "
                      f"FP64, {threads} OpenMP threads over k ({t:.1f} s)")
    if threads > 1:
        set_threads = getattr(oracle_py.lib(), "omp_set_num_threads", None)
        if set_threads is not None:
            set_threads(1)
            r1, t1 = timed(0.3 * budget_s)
            set_threads(threads)
            out["value_1core"] = round(cells * r1 / t1 / 1e9, 5)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--rhs-reps", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rhs-only", action="store_true",
                    help="only Formfunction_2 calls and a calibration copy (for counter runs)")
    ap.add_argument("--calib-reps", type=int, default=5)
    ap.add_argument("--no-implicit", action="store_true", help="skip the Newton-Krylov leg")
    ap.add_argument("--no-dynamic", action="store_true", help="skip the dynamic Smagorinsky leg")
    ap.add_argument("--no-actuator", action="store_true", help="skip the actuator-disk leg")
    ap.add_argument("--no-ib", action="store_true", help="skip the CURVIB reconstruction leg")
    ap.add_argument("--no-turbine-share", action="store_true",
                    help="skip the 768x384x48 open-boundary leg (one GPU's share of the turbine-wake case)")
    ap.add_argument("--share-only", action="store_true",
                    help="only the 768x384x48 open-boundary leg (--steps steps after --warmup), for kernel traces")
    ap.add_argument("--c5-steps", type=int, default=0,
                    help="also run the full 768x384x384 turbine-wake case for this many steps (about 200 GB of HBM)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the multi-GPU k-slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    if args.share_only:
        log("[bench] turbine-wake share only")
        print(json.dumps({"turbine_share": turbine_share_leg(args.steps, args.warmup)}), flush=True)
        return
    cfg, spec, ctx, ucat, ucont, nunk = setup(args.config, poisson=not args.rhs_only)
    Ni, Nj, Nk = spec["N"]
    cells = Ni * Nj * Nk
    log(f"[bench] {args.config} {Ni}x{Nj}x{Nk} cells={cells} unknowns={nunk}")
    s = State(ctx, ucat, ucont)
    if args.rhs_only:
        rhs_only(ctx, cfg, s, args.rhs_reps, args.calib_reps)
        return

    res = {}
    for mode, prec in (("fp64", 0), ("mixed", 1)):
        st = State(ctx, ucat, ucont)
        for _ in range(args.warmup):
            fractional_step(ctx, cfg, st, prec)
        ctx.sync()
        timers = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fractional_step(ctx, cfg, st, prec, timers=timers)
        ctx.sync()
        t1 = time.perf_counter()
        res[mode] = dict(step_ms=(t1 - t0) / args.steps * 1e3,
                         poisson_s=float(np.median([x["poisson_s"] for x in timers])),
                         iters=int(np.median([x["iters"] for x in timers])),
                         relres=max(x["relres"] for x in timers),
                         maxdiv=max(x["maxdiv"] for x in timers))
        log(f"[bench] {mode}: {res[mode]}")

    acc = accuracy_leg(ctx, s)
    log(f"[bench] mixed vs FP64 after the first projection of the seeded field: {acc}")

    ctx.contra2cart(s.uc, s.u)
    if cfg["les"]:
        from vfsgpu import workloads as W
        ctx.eddy_viscosity(s.u, W.CS_LES, s.nut)
    rhs_ms = time_rhs(ctx, cfg, s, args.rhs_reps)
    gcells = cells / (rhs_ms * 1e-3) / 1e9
    bytes_per_cell = RHS_BYTES_PER_CELL[cfg["les"]]
    achieved = bytes_per_cell * cells / (rhs_ms * 1e-3) / 1e9
    log(f"[bench] Formfunction_2 {rhs_ms:.3f} ms -> {gcells:.2f} Gcell/s, {achieved:.0f} GB/s")
    setup_ms = time_poisson_setup(ctx, 3)
    log(f"[bench] Poisson setup {setup_ms} ms")

    implicit = None
    if not args.no_implicit:
        implicit = implicit_leg(ctx, cfg, spec, s)
        log(f"[bench] implicit (CFL 2, exact JVP): {implicit}")

    dyn = None
    if not args.no_dynamic:
        dyn = dynamic_les_leg(ctx, cfg, s, 5, 0.0 if args.no_cpu else 5.0)
        log(f"[bench] dynamic Smagorinsky: {dyn}")

    act = None
    if not args.no_actuator:
        act = actuator_leg(ctx, cfg, spec, 5, not args.no_cpu)
        log(f"[bench] actuator disk: {act}")

    ib = None
    if not args.no_ib and cfg["immersed"]:
        ib = ib_leg(ctx, cfg, spec, s, 5, not args.no_cpu)
        log(f"[bench] CURVIB reconstruction: {ib}")

    share = None
    if not args.no_turbine_share and args.config == "C2":
        del s, ctx
        share = turbine_share_leg(2, 1)
        log(f"[bench] turbine-wake share: {share}")

    c5 = None
    if args.c5_steps > 0:
        if share is None:
            del s, ctx
        c5 = turbine_c5_leg(args.c5_steps, 1)
        log(f"[bench] turbine wake C5: {c5}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.config, args.cpu_budget)
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": "RHS Gcell/s + Poisson time-to-1e-8 residual, FP64 vs mixed FP32, 1 GPU",
        "value": round(gcells, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["fp64"]["step_ms"], 3),
        "ms_per_step_mixed": round(res["mixed"]["step_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config} {Ni}x{Nj}x{Nk} curvilinear channel"
                               + (" LES + immersed sphere" if cfg["les"] else ""),
                   "cells": cells, "poisson_unknowns": nunk,
                   "step": "LES + Pressure_Gradient + RK4 (4x Formfunction_2) + VolumeFlux + Poisson to 1e-8 "
                           "+ UpdatePressure + Projection + Contra2Cart + Divergence"},
        "rhs_ms": round(rhs_ms, 4),
        "poisson": {"fp64_s_to_1e-8": round(res["fp64"]["poisson_s"], 5),
                    "mixed_s_to_1e-8": round(res["mixed"]["poisson_s"], 5),
                    "fp64_iters": res["fp64"]["iters"], "mixed_iters": res["mixed"]["iters"],
                    "max_relres": max(res["fp64"]["relres"], res["mixed"]["relres"]),
                    "max_div_after_projection": max(res["fp64"]["maxdiv"], res["mixed"]["maxdiv"]),
                    "phi_rel_diff_mixed_vs_fp64": acc["phi_rel_diff_l2"],
                    "setup_ms": setup_ms},
        "accuracy_mixed_vs_fp64": acc,
        "step_ms": {"fp64": round(res["fp64"]["step_ms"], 3), "mixed": round(res["mixed"]["step_ms"], 3),
                    "speedup_mixed_vs_fp64": round(res["fp64"]["step_ms"] / res["mixed"]["step_ms"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(args.config),
                     "kernel": "Formfunction_2", "bytes_per_cell": bytes_per_cell},
        "implicit_cfl2": implicit,
        "dynamic_les": dyn,
        "actuator_disk": act,
        "ib_reconstruction": ib,
        "turbine_share": share,
        "turbine_c5": c5,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
