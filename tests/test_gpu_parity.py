"""GPU parity: every C-ABI entry point against the CPU oracle on the same
seeded inputs, at reduced sizes of the C1 (laminar channel) and C2
(stretched LES channel with an immersed sphere) generators.

Bar: FP64 results within 1e-12 relative (max |gpu - oracle| / max |oracle|);
integer-valued outputs (nvert masks, unknown numbering, masked faces) equal.
"""
import sys

import numpy as np
import pytest

from cases import gpu_case, oracle_case
from conftest import rel_err
from vfsgpu import workloads as W
import vfsgpu as V

pytestmark = pytest.mark.gpu
TOL = 1e-12
CASES = ["C1", "C2", "C1w", "C2w", "C1o", "C2o", "C1wo"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=CASES)
def case(request):
    name = request.param
    say(f"[case] {name} N={__import__('cases').SMALL[name]}")
    cfg, spec, coor, o, ucat, ucont = oracle_case(name)
    ctx = gpu_case(cfg, spec, coor)
    return dict(name=name, cfg=cfg, spec=spec, o=o, ctx=ctx, ucat=ucat, ucont=ucont)


def test_metrics_match(case):
    o, ctx = case["o"], case["ctx"]
    for n in V.METRIC_NAMES:
        g = ctx.metric(n).host()
        assert rel_err(g, o.g[n]) <= 1e-15, n


def test_nvert_bit_exact(case):
    o, ctx = case["o"], case["ctx"]
    assert np.array_equal(ctx.nvert().host(), o.nvert)


def test_contra2cart(case):
    o, ctx = case["o"], case["ctx"]
    uc, u = case["ucont"].copy(), case["ucat"].copy()
    duc, du = V.DeviceArray.from_host(uc), V.DeviceArray.from_host(u)
    o.contra2cart(uc, u)
    ctx.contra2cart(duc, du)
    assert rel_err(du.host(), u) < TOL
    assert rel_err(duc.host(), uc) < TOL


def _prepared(case):
    o, ctx = case["o"], case["ctx"]
    uc, u = case["ucont"].copy(), case["ucat"].copy()
    o.contra2cart(uc, u)
    nut = o.sca()
    if case["cfg"]["les"]:
        o.eddy_viscosity(u, W.CS_LES, nut)
    return uc, u, nut


def test_eddy_viscosity(case):
    if not case["cfg"]["les"]:
        pytest.skip("LES off in this configuration")
    o, ctx = case["o"], case["ctx"]
    uc, u, nut = _prepared(case)
    dn = ctx.sca()
    ctx.eddy_viscosity(V.DeviceArray.from_host(u), W.CS_LES, dn)
    assert rel_err(dn.host(), nut) < TOL


def test_formfunction2(case):
    o, ctx = case["o"], case["ctx"]
    uc, u, nut = _prepared(case)
    rhs0 = W.random_vector(case["cfg"], 0.3, 77)
    r = rhs0.copy()
    uc_o = uc.copy()
    o.formfunction2(uc_o, u, nut, r, 0.5)
    duc, du, dn, dr = (V.DeviceArray.from_host(x) for x in (uc, u, nut, rhs0))
    ctx.formfunction2(duc, du, dn, dr, 0.5)
    g = dr.host()
    assert rel_err(g, r) < TOL
    # masked faces are exactly zero in both
    assert np.array_equal(g == 0, r == 0)
    assert rel_err(duc.host(), uc_o) < TOL


def test_formfunction2_reference_passes(case, monkeypatch):
    """The three-pass (face / cell / projection) path agrees too."""
    monkeypatch.setenv("VFS_FF2_PATH", "passes")
    test_formfunction2(case)


@pytest.mark.parametrize("scheme", [dict(weno=1), dict(inviscid=1), dict(inviscid=1, second_order=1), dict(skew=1),
                                    dict(clark=1), dict(skew=1, clark=1, second_order=1), dict(skew=1, inviscid=1)],
                         ids=["weno", "inviscid", "inviscid-2nd", "skew", "clark", "skew-clark-2nd", "skew-inviscid"])
def test_formfunction2_weno(case, scheme):
    """-weno 1: WENO3 convective face values (momentum.c:782-786 and the j/k
    counterparts); -inv: WENO3 whatever else is set, the 2nd-order divergence
    and no viscous term (momentum.c:757-773, 1582, 1621); -skew: half the
    divergence form plus the advective face fluxes interpolated to the cells
    (momentum.c:798-811, 1596-1608), whose cell part -inv skips; -clark: the
    Clark model stress in the viscous face fluxes (momentum.c:907-925); all
    on the same fields."""
    cfg = dict(case["cfg"], **scheme)
    _, spec, coor, o, ucat, ucont = oracle_case(case["name"])
    o = __import__("oracle_py").Oracle(cfg, coor)
    if spec["immersed"]:
        o.nvert_sphere(*W.SPHERE)
    o.wall_mask()
    uc, u = ucont.copy(), ucat.copy()
    o.contra2cart(uc, u)
    nut = o.sca()
    if cfg["les"]:
        o.eddy_viscosity(u, W.CS_LES, nut)
    ctx = gpu_case(cfg, spec, coor)
    duc, du, dn = (V.DeviceArray.from_host(x) for x in (uc, u, nut))
    rhs = o.vec()
    o.formfunction2(uc, u, nut, rhs, 1.0)
    drhs = ctx.vec()
    ctx.formfunction2(duc, du, dn, drhs, 1.0)
    assert rel_err(drhs.host(), rhs) < TOL
    ctx.close()


def test_derivatives_reject_skew_and_clark(case):
    """The JVP / VJP do not differentiate -skew or -clark: such contexts are
    refused before any work is enqueued."""
    _, spec, coor, o, ucat, ucont = oracle_case(case["name"])
    for opt in ("skew", "clark"):
        ctx = gpu_case(dict(case["cfg"], **{opt: 1}), spec, coor)
        U, u, w, jw, nut = ctx.vec(), ctx.vec(), ctx.vec(), ctx.vec(), ctx.sca()
        for f in (ctx.rhs_jvp, ctx.rhs_vjp):
            with pytest.raises(V.VfsError) as e:
                f(U, u, nut, w, jw)
            assert e.value.rc == -2
            # wall-function walls are refused first, for their own reason
            assert opt in str(e.value) or -1 in case["cfg"]["bctype"]
        ctx.close()


def test_pressure_gradient(case):
    o, ctx = case["o"], case["ctx"]
    p = W.random_scalar(case["cfg"], 1.0, 555)
    dp = o.vec()
    pp = p.copy()
    o.pressure_gradient(pp, dp)
    dpp, ddp = V.DeviceArray.from_host(p), ctx.vec()
    ctx.pressure_gradient(dpp, ddp)
    assert rel_err(ddp.host(), dp) < TOL
    assert rel_err(dpp.host(), pp) == 0.0


def test_ib_bc(case):
    o, ctx = case["o"], case["ctx"]
    uc, u, nut = _prepared(case)
    uc2, u2 = uc.copy(), u.copy()
    o.ib_bc(uc2, u2)
    duc, du = V.DeviceArray.from_host(uc), V.DeviceArray.from_host(u)
    ctx.ib_bc(duc, du)
    assert rel_err(duc.host(), uc2) < TOL
    assert rel_err(du.host(), u2) < TOL


def test_formfunction_snes(case):
    o, ctx = case["o"], case["ctx"]
    cfg = case["cfg"]
    uc, u, nut = _prepared(case)
    U = uc + 0.01 * W.random_vector(cfg, 1.0, 31)
    Uo = uc.copy()
    rhs_o = W.random_vector(cfg, 0.1, 32)
    dP = W.random_vector(cfg, 0.1, 33)
    lu, uu, r = o.vec(), u.copy(), o.vec()
    o.formfunction_snes(U, Uo, rhs_o, dP, nut, lu, uu, r)
    d = [V.DeviceArray.from_host(x) for x in (U, Uo, rhs_o, dP, nut)]
    dlu, duu, dr = ctx.vec(), V.DeviceArray.from_host(u), ctx.vec()
    ctx.formfunction_snes(*d, dlu, duu, dr)
    assert rel_err(dr.host(), r) < TOL
    assert rel_err(dlu.host(), lu) < TOL


def test_rk4_step(case):
    """The four RungeKutta stages (rhs.c:1586, 1637-1659) chained."""
    o, ctx = case["o"], case["ctx"]
    cfg = case["cfg"]
    uc, u, nut = _prepared(case)
    dP = W.random_vector(cfg, 0.01, 34)
    u0 = uc.copy()
    a_uc, a_u, rhs = uc.copy(), u.copy(), o.vec()
    du0, duc, du, dn, ddp, drhs = (V.DeviceArray.from_host(x) for x in (u0, uc, u, nut, dP, rhs))
    for alfa in (0.25, 1 / 3.0, 0.5, 1.0):
        o.rk_stage(u0, a_uc, a_u, nut, dP, alfa, rhs)
        ctx.rk_stage(du0, duc, du, dn, ddp, alfa, drhs)
    assert rel_err(duc.host(), a_uc) < TOL
    assert rel_err(du.host(), a_u) < TOL


def test_min_dt(case):
    o, ctx = case["o"], case["ctx"]
    uc = case["ucont"] + 0.3
    assert rel_err(ctx.min_dt(V.DeviceArray.from_host(uc)), o.min_dt(uc)) < TOL


# ---------------------------------------------------------------------------
# pressure projection
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def poisson(case):
    o, ctx = case["o"], case["ctx"]
    n = ctx.poisson_setup()
    o.setup_gid()
    o.poisson_lhs()
    return n


def test_poisson_numbering_and_coefficients(case, poisson):
    o, ctx = case["o"], case["ctx"]
    assert poisson == o.n
    assert np.array_equal(ctx.poisson_gid().host(), o.gid)
    coef = ctx.poisson_coefficients().host().reshape(19, -1).T
    ref = o.poisson_stencil().reshape(-1, 19)
    # the oracle stencil is pre-drop; apply the drop rule of poisson.c:1604-1608
    rp, col, val = o.csr
    dense = np.zeros_like(ref)
    gid = o.gid
    P = o.P
    nb = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                   [1, -1, 0], [-1, 1, 0], [-1, -1, 0], [0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1], [1, 0, 1],
                   [1, 0, -1], [-1, 0, 1], [-1, 0, -1]])
    G = gid.reshape(o.mz, o.my, o.mx)
    interior = np.zeros_like(G, dtype=bool)
    interior[1:-1, 1:-1, 1:-1] = True
    cells = np.nonzero((G >= 0) & interior)[0] if False else np.nonzero(((G >= 0) & interior).ravel())[0]
    for t in cells[:: max(1, len(cells) // 400)]:
        i, j, k = t % o.mx, (t // o.mx) % o.my, t // (o.mx * o.my)
        row = int(gid[t])
        for m in range(19):
            q = ((k + nb[m, 2]) * o.my + j + nb[m, 1]) * o.mx + i + nb[m, 0]
            c = int(gid[q])
            hit = [val[s] for s in range(rp[row], rp[row + 1]) if col[s] == c]
            expect = hit[0] if (hit and (m == 0 or c != row)) else 0.0
            assert coef[t, m] == expect, (t, m)


def test_poisson_apply(case, poisson):
    o, ctx = case["o"], case["ctx"]
    x = W.random_scalar(case["cfg"], 1.0, 901)
    xp = x.copy()
    # periodic ghosts of x as the reference's lPhi would hold them
    X = xp.reshape(o.mz, o.my, o.mx)
    X[:, :, 0] = X[:, :, -2]; X[:, :, -1] = X[:, :, 1]; X[0] = X[-2]; X[-1] = X[1]
    y = o.sca()
    o.csr_apply(xp, y)
    dy = ctx.sca()
    ctx.poisson_apply(V.DeviceArray.from_host(xp), dy)
    assert rel_err(dy.host(), y) < TOL


def test_volume_flux_and_rhs(case, poisson):
    o, ctx = case["o"], case["ctx"]
    uc = case["ucont"].copy()
    f_o, a_o = o.volume_flux(uc)   # the plain variant (poisson.c:3306-3500) when no body is immersed
    duc = V.DeviceArray.from_host(case["ucont"])
    f_g, a_g = ctx.volume_flux(duc)
    assert rel_err(duc.host(), uc) < TOL
    assert abs(f_g - f_o) <= 1e-12 * max(1.0, a_o) and rel_err(a_g, a_o) < TOL
    b = o.sca()
    o.poisson_rhs(uc, b)
    db = ctx.sca()
    ctx.poisson_rhs(duc, db)
    assert rel_err(db.host(), b) < TOL


def test_volume_flux_plain_variant(case):
    """VolumeFlux with immersed == 0 and a marked region (poisson.c:3306-3500)."""
    cfg = dict(case["cfg"], immersed=0)
    _, spec, coor, _, _, _ = oracle_case(case["name"])
    o = __import__("oracle_py").Oracle(cfg, coor)
    nv = np.zeros((o.mz, o.my, o.mx))
    nv[3:6, 2:5, 4:7] = 1.0
    o.nvert[:] = nv.ravel()
    o.wall_mask()   # the library keeps the wall-function cells marked
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(nv.ravel()))
    uc = case["ucont"].copy()
    f_o, a_o = o.volume_flux(uc)
    duc = V.DeviceArray.from_host(case["ucont"])
    f_g, a_g = ctx.volume_flux(duc)
    assert np.array_equal(duc.host(), uc)
    assert f_g == 0.0 and f_o == 0.0 and rel_err(a_g, a_o) < TOL
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_poisson_solve_converged_solution(case, poisson, precision):
    """Parity on the converged solution: both sides reach their residual and
    the GPU phi matches the oracle's 1e-13 solution to the conditioning."""
    o, ctx = case["o"], case["ctx"]
    uc = case["ucont"].copy()
    o.volume_flux(uc)
    b = o.sca()
    o.poisson_rhs(uc, b)
    phi_o = o.sca()
    it_o, rel_o = o.poisson_solve(b, phi_o, rtol=1e-13)
    dphi = ctx.sca()
    it, rel = ctx.poisson_solve(V.DeviceArray.from_host(b), dphi, rtol=1e-10, precision=precision)
    say(f"  solve precision={precision} iters={it} relres={rel:.3e} (oracle {it_o} its, {rel_o:.1e})")
    assert rel <= 1e-10
    assert rel_err(dphi.host(), phi_o) < 1e-7


@pytest.mark.parametrize("block", [False, True])
def test_poisson_single_level_dense_inverse(block):
    """A grid of at most 1024 unknowns keeps one multigrid level, which the
    preconditioner solves with the dense inverse built at set-up (Gauss-Jordan
    with partial pivoting on the device; with the marked block, 27 cells are
    not unknowns).  With that exact preconditioner FP64 GMRES converges in one
    iteration (two at most); a wrong inverse would need many."""
    say(f"[case] single-level C1 N=(10, 8, 9) block={block}")
    cfg, spec, coor, o, ucat, ucont = oracle_case("C1", N=(10, 8, 9))
    ctx = gpu_case(cfg, spec, coor)
    if block:
        nv = np.zeros((o.mz, o.my, o.mx))
        nv[3:6, 2:5, 4:7] = 1.0
        o.nvert[:] = nv.ravel()
        o.wall_mask()
        ctx.set_nvert(V.DeviceArray.from_host(nv.ravel()))
    n = ctx.poisson_setup()
    o.setup_gid()
    o.poisson_lhs()
    assert n == o.n and n <= 1024
    uc = ucont.copy()
    o.volume_flux(uc)
    b = o.sca()
    o.poisson_rhs(uc, b)
    phi_o = o.sca()
    o.poisson_solve(b, phi_o, rtol=1e-13)
    for precision, max_it, tol in ((0, 2, 1e-12), (1, 10, 1e-10)):
        dphi = ctx.sca()
        it, rel = ctx.poisson_solve(V.DeviceArray.from_host(b), dphi, rtol=tol, precision=precision)
        say(f"  precision={precision} iters={it} relres={rel:.2e}")
        assert it <= max_it and rel <= tol
        assert rel_err(dphi.host(), phi_o) < 1e-7
    ctx.close()


def test_projection_update_divergence(case, poisson):
    o, ctx = case["o"], case["ctx"]
    cfg = case["cfg"]
    phi = W.random_scalar(cfg, 1e-3, 902)
    P = phi.reshape(o.mz, o.my, o.mx)
    P[:, :, 0] = P[:, :, -2]; P[:, :, -1] = P[:, :, 1]; P[0] = P[-2]; P[-1] = P[1]
    uc = case["ucont"].copy()
    o.projection(phi, uc)
    duc = V.DeviceArray.from_host(case["ucont"])
    ctx.projection(V.DeviceArray.from_host(phi), duc)
    assert rel_err(duc.host(), uc) < TOL
    p = W.random_scalar(cfg, 1.0, 903)
    pp, ph = p.copy(), phi.copy()
    o.update_pressure(pp, ph)
    dp, dph = V.DeviceArray.from_host(p), V.DeviceArray.from_host(phi)
    ctx.update_pressure(dp, dph)
    assert rel_err(dp.host(), pp) < TOL
    m_o, div_o = o.divergence(uc)
    ddiv = ctx.sca()
    m_g = ctx.divergence(duc, ddiv)
    assert rel_err(ddiv.host(), div_o) < TOL and rel_err(m_g, m_o) < TOL


def test_full_fractional_step(case, poisson):
    """LES -> Pressure_Gradient -> RK4 momentum -> VolumeFlux -> Poisson
    (1e-10) -> UpdatePressure -> Projection -> Contra2Cart -> Divergence,
    GPU against oracle."""
    o, ctx = case["o"], case["ctx"]
    cfg = case["cfg"]
    uc, u, _ = _prepared(case)
    p = W.random_scalar(cfg, 0.01, 904)
    # oracle
    nut = o.sca()
    if cfg["les"]:
        o.eddy_viscosity(u, W.CS_LES, nut)
    pp, dP = p.copy(), o.vec()
    o.pressure_gradient(pp, dP)
    u0, a_uc, a_u, rhs = uc.copy(), uc.copy(), u.copy(), o.vec()
    for alfa in (0.25, 1 / 3.0, 0.5, 1.0):
        o.rk_stage(u0, a_uc, a_u, nut, dP, alfa, rhs)
    o.volume_flux(a_uc)
    b = o.sca()
    o.poisson_rhs(a_uc, b)
    phi = o.sca()
    o.poisson_solve(b, phi, rtol=1e-13)
    o.update_pressure(pp, phi)
    o.projection(phi, a_uc)
    o.contra2cart(a_uc, a_u)
    md_o, _ = o.divergence(a_uc)
    # gpu
    D = {k: V.DeviceArray.from_host(v) for k, v in dict(uc=uc, u=u, p=p).items()}
    dn, ddp, drhs, dphi, db = ctx.sca(), ctx.vec(), ctx.vec(), ctx.sca(), ctx.sca()
    if cfg["les"]:
        ctx.eddy_viscosity(D["u"], W.CS_LES, dn)
    ctx.pressure_gradient(D["p"], ddp)
    du0 = V.DeviceArray.from_host(uc)
    for alfa in (0.25, 1 / 3.0, 0.5, 1.0):
        ctx.rk_stage(du0, D["uc"], D["u"], dn, ddp, alfa, drhs)
    ctx.volume_flux(D["uc"])
    ctx.poisson_rhs(D["uc"], db)
    ctx.poisson_solve(db, dphi, rtol=1e-11, precision=0)
    ctx.update_pressure(D["p"], dphi)
    ctx.projection(dphi, D["uc"])
    ctx.contra2cart(D["uc"], D["u"])
    md = ctx.divergence(D["uc"])
    say(f"  step maxdiv gpu {md:.3e} oracle {md_o:.3e}")
    assert rel_err(D["uc"].host(), a_uc) < 1e-9
    assert rel_err(D["p"].host(), pp) < 1e-7
    assert md < 1e-8
