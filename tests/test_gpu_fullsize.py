"""Parity at the benchmark sizes: C1 (64^3 laminar channel) and C2/C4
(256x128x128 stretched LES channel with the immersed sphere).

The single-pass routines are compared with the CPU oracle on the full grids
(each oracle call takes about a second here).  The Poisson solve, whose
oracle (GMRES + Jacobi) would take minutes at this size, is checked through
the properties that do not depend on size: the true FP64 residual, the
agreement of the FP64 and mixed-precision solutions, and the divergence
left after the projection.  The derivatives are checked by the adjoint
identity and JVP linearity at C4 size.
"""
import numpy as np
import pytest

from cases import gpu_case, oracle_case
from conftest import rel_err
from vfsgpu import workloads as W
import vfsgpu as V

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", params=["C1", "C2"])
def big(request):
    name = request.param
    N = W.BASE[name]["N"]
    say(f"[fullsize] {name} {N}")
    cfg, spec, coor, o, ucat, ucont = oracle_case(name, N)
    ctx = gpu_case(cfg, spec, coor)
    uc, u = ucont.copy(), ucat.copy()
    o.contra2cart(uc, u)
    nut = o.sca()
    if cfg["les"]:
        o.eddy_viscosity(u, W.CS_LES, nut)
    yield dict(name=name, cfg=cfg, o=o, ctx=ctx, ucont=ucont, ucat=ucat, uc=uc, u=u, nut=nut)
    ctx.close()


def test_metrics_and_mask(big):
    o, ctx = big["o"], big["ctx"]
    for n in ("csi", "aj", "kzet", "jaj"):
        assert rel_err(ctx.metric(n).host(), o.g[n]) <= 1e-15, n
    assert np.array_equal(ctx.nvert().host(), o.nvert)


def test_contra2cart_les_formfunction2(big):
    o, ctx, cfg = big["o"], big["ctx"], big["cfg"]
    duc, du = V.DeviceArray.from_host(big["ucont"]), V.DeviceArray.from_host(big["ucat"])
    ctx.contra2cart(duc, du)
    assert rel_err(du.host(), big["u"]) < TOL
    dn = ctx.sca()
    if cfg["les"]:
        ctx.eddy_viscosity(du, W.CS_LES, dn)
        assert rel_err(dn.host(), big["nut"]) < TOL
    rhs = o.vec()
    uc = big["uc"].copy()
    o.formfunction2(uc, big["u"], big["nut"], rhs, 1.0)
    drhs = ctx.vec()
    ctx.formfunction2(duc, du, dn, drhs, 1.0)
    err = rel_err(drhs.host(), rhs)
    say(f"  Formfunction_2 rel err {err:.2e}")
    assert err < TOL


def test_pressure_gradient_projection_divergence(big):
    o, ctx, cfg = big["o"], big["ctx"], big["cfg"]
    p = W.random_scalar(cfg, 1.0, 811)
    dp = o.vec()
    o.pressure_gradient(p.copy(), dp)
    ddp = ctx.vec()
    ctx.pressure_gradient(V.DeviceArray.from_host(p), ddp)
    assert rel_err(ddp.host(), dp) < TOL
    phi = W.random_scalar(cfg, 1e-3, 812)
    P = phi.reshape(o.mz, o.my, o.mx)
    P[:, :, 0] = P[:, :, -2]; P[:, :, -1] = P[:, :, 1]; P[0] = P[-2]; P[-1] = P[1]
    uc = big["uc"].copy()
    o.projection(phi, uc)
    duc = V.DeviceArray.from_host(big["uc"])
    ctx.projection(V.DeviceArray.from_host(phi), duc)
    assert rel_err(duc.host(), uc) < TOL
    m_o, _ = o.divergence(uc)
    assert rel_err(ctx.divergence(duc), m_o) < TOL


def test_poisson_full_size(big):
    """RHS parity with the oracle, then FP64 and mixed solves to 1e-10: true
    residual, agreement of the two solutions, divergence after projection."""
    o, ctx = big["o"], big["ctx"]
    ctx.poisson_setup()
    o.setup_gid()
    uc = big["uc"].copy()
    if big["cfg"]["immersed"]:
        o.volume_flux(uc)
    b = o.sca()
    o.poisson_rhs(uc, b)
    duc = V.DeviceArray.from_host(big["uc"])
    ctx.volume_flux(duc)
    db = ctx.sca()
    ctx.poisson_rhs(duc, db)
    assert rel_err(db.host(), b) < TOL
    sols = {}
    for prec in (0, 1):
        phi = ctx.sca()
        it, rel = ctx.poisson_solve(db, phi, rtol=1e-10, precision=prec, maxit=1000)
        say(f"  precision {prec}: {it} iterations, rel res {rel:.2e}")
        assert rel <= 1e-10
        # measured with the over-corrected W-cycle: 22 / 22 (C1), 26 / 28 (C2);
        # a weaker hierarchy or cycle shows here before any residual test fails
        assert it <= {"C1": 30, "C2": 36}[big["name"]]
        sols[prec] = phi
    a, m = sols[0].host(), sols[1].host()
    diff = np.linalg.norm(m - a) / np.linalg.norm(a)
    say(f"  |phi_mixed - phi_fp64| / |phi_fp64| = {diff:.2e}")
    # both within 1e-10 relative residual: A (phi_m - phi_a) = r_a - r_m is at
    # most 2e-10 |b|; the error this leaves in phi is far below the bound
    # (measured about 1e-11; 3e-9 for solves to 1e-8)
    assert diff < 1e-8
    us = []
    for prec in (0, 1):
        u2 = V.DeviceArray.from_host(duc.host())
        ctx.projection(sols[prec], u2)
        us.append(u2)
        maxdiv = ctx.divergence(u2)
        m0 = ctx.divergence(duc)
        say(f"  precision {prec}: max|div U| before {m0:.3e} after {maxdiv:.3e}")
        # |r|_inf <= |r|_2 <= 1e-10 |b|_2 <= 1e-10 sqrt(n) |b|_inf
        assert maxdiv < 1e-5 * m0
    du = np.linalg.norm(us[1].host() - us[0].host()) / np.linalg.norm(us[0].host())
    say(f"  |U_mixed - U_fp64| / |U_fp64| after projection = {du:.2e}")
    assert du <= 1e-9


def test_derivatives_full_size(big):
    """Adjoint identity and linearity of the JVP at C4 size (seeds 14001/24001)."""
    if big["name"] != "C2":
        pytest.skip("C4 is the C2 grid")
    ctx, cfg = big["ctx"], big["cfg"]
    U = V.DeviceArray.from_host(big["uc"])
    du = V.DeviceArray.from_host(big["u"])
    dn = V.DeviceArray.from_host(big["nut"])
    w = W.unit_vector(cfg, 14001)
    v = W.unit_vector(cfg, 24001)
    jw, jtv = ctx.vec(), ctx.vec()
    ctx.rhs_jvp(U, du, dn, V.DeviceArray.from_host(w), jw)
    ctx.rhs_vjp(U, du, dn, V.DeviceArray.from_host(v), jtv)
    Jw, JTv = jw.host(), jtv.host()
    lhs, rhs = float(v @ Jw), float(JTv @ w)
    scale = np.linalg.norm(v) * np.linalg.norm(Jw)
    say(f"  adjoint identity rel {abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= 1e-12 * scale
    j2 = ctx.vec()
    ctx.rhs_jvp(U, du, dn, V.DeviceArray.from_host(2.0 * w), j2)
    assert rel_err(j2.host(), 2.0 * Jw) < 1e-14


def test_dynamic_smagorinsky_full_size(big):
    """The dynamic procedure (-les 2) and its eddy viscosity on the full grid
    against the oracle (about 20 s of oracle time at C2)."""
    o, ctx = big["o"], big["ctx"]
    du = V.DeviceArray.from_host(big["u"])
    cs, nut = ctx.sca(), ctx.sca()
    ctx.smagorinsky_constant(du, cs)
    ctx.eddy_viscosity_cs(du, cs, nut)
    say("  dynamic Cs on the GPU done; oracle ...")
    want = o.smagorinsky_constant(big["u"])
    nut_o = o.sca()
    o.eddy_viscosity_cs(big["u"], want, nut_o)
    got = cs.host()
    say(f"  Cs rel err {rel_err(got, want):.2e}, max Cs {want.max():.3f}")
    assert rel_err(got, want) < TOL
    assert np.array_equal(got == 0, want == 0)
    assert rel_err(nut.host(), nut_o) < TOL
