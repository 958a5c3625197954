"""The explicit fractional step in the reference's order: RungeKutta with its
boundary work (rhs.c:1578-1700) and Flow_Solver around it
(solvers.c:590-1390).

CPU tests pin the oracle's composition:
  - on a periodic body-free case vo_runge_kutta is U0 = U and the four
    stages of vo_rk_stage, bit for bit;
  - with a k inlet and outlet the final FormBCS (outflow_scale 1) leaves an
    outflow equal to the inflow, while the stages alone do not.
GPU tests run vfs_runge_kutta and two whole steps of vfsgpu.step.FlowSolver
against the oracle run in the same order, on C1wo (wall-function bottom, k
inlet and outlet), C2o (stretched wavy grid, LES, a triangulated sphere with
the CURVIB reconstruction, k inlet and outlet) and C2 with the same body.
"""
import sys

import numpy as np
import pytest

import oracle_py as OP
from cases import oracle_case
from conftest import rel_err
from oracle_py import IbBody
from vfsgpu import workloads as W
import vfsgpu as V
from vfsgpu.step import FlowSolver

N_SMALL = {"C1wo": (16, 12, 14), "C2o": (32, 24, 20), "C2": (32, 24, 20), "C1": (16, 12, 14)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(name):
    cfg, spec, coor, o, ucat, ucont = oracle_case(name, N_SMALL[name])
    body = None
    if spec["immersed"]:
        x, e, n = W.icosphere(*W.SPHERE, level=3)
        body = (x, e, n)
        o.nvert_body(IbBody(x, e, n))
    return cfg, spec, coor, o, ucat, ucont, body


def _oracle_step(o, cfg, coor, s, body, nodes, first, inlet_flux=-1.0):
    """Flow_Solver's explicit step with the oracle (solvers.c:599-1385)."""
    def reconstruct():
        if cfg["immersed"] and body is not None:
            o.ib_interpolate(body, nodes, s["uc"], s["u"], s["p"], s["ustar"])
    if first:
        reconstruct()
        o.ib_bc(s["uc"], s["u"])
    kflux = o.calc_k_flux(s["uc"]) if cfg["k_periodic"] and not cfg["dpdz_set"] else None
    if cfg["les"]:
        o.contra2cart(s["uc"], s["u"])
        o.eddy_viscosity(s["u"], W.CS_LES, s["nut"])
    dp = o.vec()
    o.pressure_gradient(s["p"], dp, kflux=kflux, inlet_flux=inlet_flux)
    inlet = not cfg["k_periodic"] and cfg["bctype"][4] == 5
    u0, rhs, fi, fo = o.runge_kutta(s["uc"], s["u"], s["nut"], dp, coor=coor,
                                    inflow=OP.VoInflowOpt(1, 0, -1.0, 0.0) if inlet else None,
                                    ubcs=s.get("ubcs"), body=body if cfg["immersed"] else None,
                                    nodes=nodes if cfg["immersed"] else None, p=s["p"], ustar=s["ustar"])
    o.volume_flux(s["uc"])
    b = o.sca()
    o.poisson_rhs(s["uc"], b)
    it, rel = o.poisson_solve(b, s["phi"], rtol=1e-13)
    o.update_pressure(s["p"], s["phi"])
    o.projection(s["phi"], s["uc"])
    reconstruct()
    m, _ = o.divergence(s["uc"])
    o.ib_bc(s["uc"], s["u"])
    o.contra2cart(s["uc"], s["u"])
    return dict(flux_in=fi, flux_out=fo, maxdiv=m, relres=rel)


# ---------------------------------------------------------------------------
# CPU: the oracle's composition
# ---------------------------------------------------------------------------
def test_oracle_runge_kutta_is_four_stages_on_a_periodic_case():
    cfg, spec, coor, o, ucat, ucont, body = _case("C1")
    dp = W.random_vector(cfg, 0.01, 5)
    nut = o.sca()
    uc1, u1 = ucont.copy(), ucat.copy()
    u0, rhs, fi, fo = o.runge_kutta(uc1, u1, nut, dp)
    uc2, u2, r2 = ucont.copy(), ucat.copy(), o.vec()
    base = uc2.copy()
    for a in (0.25, 1 / 3.0, 0.5, 1.0):
        OP.lib().vo_rk_stage(OP.C.byref(o.c), OP.C.byref(o.geom), OP.ptr(o.nvert), OP.ptr(base), OP.ptr(uc2),
                             OP.ptr(u2), OP.ptr(nut), OP.ptr(dp), OP.C.c_double(a), OP.ptr(r2))
    assert np.array_equal(uc1, uc2) and np.array_equal(u1, u2) and np.array_equal(rhs, r2)
    assert np.array_equal(u0, ucont) and fi == 0 and fo == 0


def test_oracle_runge_kutta_balances_the_outflow():
    cfg, spec, coor, o, ucat, ucont, body = _case("C1wo")
    dp = o.vec()
    uc, u, ubcs = ucont.copy(), ucat.copy(), o.vec()
    u0, rhs, fi, fo = o.runge_kutta(uc, u, o.sca(), dp, coor=coor, inflow=OP.VoInflowOpt(1, 0, -1.0, 0.0), ubcs=ubcs)
    out = o.outflow_flux(uc)
    assert abs(out - fi) <= 1e-13 * abs(fi)
    # without the final correction the stages leave an imbalance
    assert abs(fo - fi) > 1e-6 * abs(fi)
    # the inflow is the uniform profile through the fluid inlet faces
    _, _, inlet = o.inlet_area()
    assert abs(fi - inlet) <= 1e-13 * inlet


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def _gpu(cfg, spec, coor, o, body):
    ctx = V.Ctx(cfg)
    dcoor = V.DeviceArray.from_host(coor)
    ctx.build_metrics(dcoor)
    surf = dn = None
    n = 0
    if body is not None:
        surf = V.IbSurface(*body)
        ctx.nvert_body(surf)
        dn, n = ctx.ib_search(surf)
    assert np.array_equal(ctx.nvert().host(), o.nvert)
    return ctx, dcoor, surf, dn, n


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["C1wo", "C2o", "C2"])
def test_gpu_runge_kutta(name):
    say(f"[step] vfs_runge_kutta {name} N={N_SMALL[name]}")
    cfg, spec, coor, o, ucat, ucont, body = _case(name)
    ctx, dcoor, surf, dn, n = _gpu(cfg, spec, coor, o, body)
    nodes = o.ib_search(IbBody(*body)) if body is not None else None
    inlet = cfg["bctype"][4] == 5 and not cfg["k_periodic"]
    nut = o.sca()
    if cfg["les"]:
        u = ucat.copy()
        o.contra2cart(ucont.copy(), u)
        o.eddy_viscosity(u, W.CS_LES, nut)
    dp = W.random_vector(cfg, 0.01, 7)
    p = W.random_scalar(cfg, 0.1, 8)
    ustar = o.sca()
    uc, u, pp, us, ubcs = ucont.copy(), ucat.copy(), p.copy(), ustar.copy(), o.vec()
    u0, rhs, fi, fo = o.runge_kutta(uc, u, nut, dp, coor=coor, inflow=OP.VoInflowOpt(1, 0, -1.0, 0.0) if inlet else None,
                                    ubcs=ubcs, body=IbBody(*body) if body else None, nodes=nodes, p=pp, ustar=us)
    d = {k: V.DeviceArray.from_host(a) for k, a in dict(uc=ucont, u=ucat, p=p, us=ustar, nut=nut, dp=dp).items()}
    d_ubcs, d_u0, d_rhs = ctx.vec(), ctx.vec(), ctx.vec()
    gfi, gfo = ctx.runge_kutta(d["uc"], d["u"], d["nut"], d["dp"], d_u0, d_rhs,
                               inflow=V.InflowOptions(1, 0, -1.0, 0.0) if inlet else None, coor=dcoor, ubcs=d_ubcs,
                               surface=surf, nodes=dn, n_nodes=n, p=d["p"], ustar=d["us"])
    for g, h, what in ((d["uc"], uc, "ucont"), (d["u"], u, "ucat"), (d_u0, u0, "u0"), (d_rhs, rhs, "rhs"),
                       (d["p"], pp, "p")):
        assert rel_err(g.host(), h) <= 1e-12, what
    assert abs(gfi - fi) <= 1e-14 * max(1.0, abs(fi)) and abs(gfo - fo) <= 1e-12 * max(1.0, abs(fo))
    if inlet:
        assert abs(ctx.outflow_flux(d["uc"]) - fi) <= 1e-12 * abs(fi)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["C1wo", "C2o"])
def test_gpu_flow_solver_two_steps(name):
    say(f"[step] FlowSolver, two steps, {name} N={N_SMALL[name]}")
    cfg, spec, coor, o, ucat, ucont, body = _case(name)
    ctx, dcoor, surf, dn, n = _gpu(cfg, spec, coor, o, body)
    nodes = o.ib_search(IbBody(*body)) if body is not None else None
    ctx.poisson_setup()
    fs = FlowSolver(ctx, V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), coor=dcoor,
                    body=(surf, dn, n) if body is not None else None, cs=W.CS_LES, precision=0, rtol=1e-12)
    s = dict(uc=ucont.copy(), u=ucat.copy(), p=o.sca(), nut=o.sca(), ustar=o.sca(), phi=o.sca(), ubcs=o.vec())
    for step in range(2):
        g = fs.step()
        h = _oracle_step(o, cfg, coor, s, IbBody(*body) if body else None, nodes, step == 0)
        say(f"[step] {name} step {step}: gpu {g} oracle {h}")
        assert g["relres"] <= 1e-12
        assert rel_err(fs.uc.host(), s["uc"]) <= 1e-9
        assert rel_err(fs.u.host(), s["u"]) <= 1e-9
        assert rel_err(fs.p.host(), s["p"]) <= 1e-8
        assert g["maxdiv"] <= 1e-9 and h["maxdiv"] <= 1e-9
        if cfg["bctype"][4] == 5:
            assert abs(g["flux_out_corrected"] - g["flux_in"]) <= 1e-12 * abs(g["flux_in"])


@pytest.mark.gpu
def test_gpu_flux_forcing():
    """Calc_k_Flux and the flux forcing of Pressure_Gradient (momentum.c:
    387-398) on the C1 channel with a sphere, k-periodic without dpdz_set,
    against the oracle; then two steps of FlowSolver with -flux 10% above
    the initial mean flux."""
    say("[step] flux forcing, C2 with dpdz_set = 0")
    cfg, spec, coor, o, ucat, ucont, body = _case("C2")
    cfg = dict(cfg, dpdz_set=0)
    o = OP.Oracle(cfg, coor)
    x, e, n = body
    o.nvert_body(IbBody(x, e, n))
    ctx, dcoor, surf, dn, nn = _gpu(cfg, spec, coor, o, body)
    duc = V.DeviceArray.from_host(ucont)
    F, A = ctx.calc_k_flux(duc)
    Fo, Ao = o.calc_k_flux(ucont.copy())
    # the device sums each plane in a fixed tree order
    assert abs(F - Fo) <= 1e-14 * np.abs(ucont).max() * cfg["mx"] * cfg["my"] and abs(A - Ao) <= 1e-14 * Ao
    fin = 1.1 * Fo
    ctx.set_inlet_flux(fin)
    p = W.random_scalar(cfg, 0.1, 9)
    dp = ctx.vec()
    ctx.pressure_gradient(V.DeviceArray.from_host(p), dp)
    dpo = o.vec()
    o.pressure_gradient(p.copy(), dpo, kflux=(Fo, Ao), inlet_flux=fin)
    assert rel_err(dp.host(), dpo) <= 1e-12
    nodes = o.ib_search(IbBody(*body))
    ctx.poisson_setup()
    fs = FlowSolver(ctx, V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), coor=dcoor,
                    body=(surf, dn, nn), cs=W.CS_LES, precision=0, rtol=1e-12, inlet_flux=fin)
    s = dict(uc=ucont.copy(), u=ucat.copy(), p=o.sca(), nut=o.sca(), ustar=o.sca(), phi=o.sca())
    for step in range(2):
        g = fs.step()
        h = _oracle_step(o, cfg, coor, s, IbBody(*body), nodes, step == 0, inlet_flux=fin)
        say(f"[step] flux-forced C2 step {step}: gpu {g} oracle {h}")
        assert rel_err(fs.uc.host(), s["uc"]) <= 1e-9
        assert rel_err(fs.p.host(), s["p"]) <= 1e-8
