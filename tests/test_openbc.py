"""Open k boundaries: Calc_Inlet_Area (bcs.c:216-328), InflowFlux at a k = 0
inlet (bcs.c:329-706), OutflowFlux at a k = mz-1 outlet (bcs.c:801-870) and
the outlet branch of FormBCS (bcs.c:2285-2418).  The device sums each plane
in a fixed tree order and the oracle in the reference's j, i loop, so sums
and what depends on them are compared to 1e-14 relative, the face values
themselves bit for bit.

CPU tests pin the oracle:
  - Contra2Cart_single (rhs.c:34-55) against a linear solve with the rows
    csi, eta, zet;
  - on a uniform grid every fluid k plane has the area Lx Ly less the
    non-fluid faces, and so has the inlet plane;
  - a uniform inflow gives U^zeta = dx dy at every fluid face, a total of
    Lx Ly, ubcs equal to the ghost velocity and a face-averaged velocity of
    (0, 0, 1); the parabola 1.5 y (2 - y) gives Lx (2 + dy^2 / 4), the
    midpoint rule's value for a quadratic;
  - after the outlet correction the outflow equals the inflow, and ubcs is
    the last cells' velocity.
GPU tests compare the HIP path with the oracle.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle_py import Oracle, lib, ptr
from vfsgpu import workloads as W
import vfsgpu as V
import ctypes as C


def open_case(N=(20, 16, 18), beta=0.0, amp=0.0, sphere=False):
    Ni, Nj, Nk = N
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=0, k_periodic=0,
               bctype=[7, 7, 1, 1, 1, 1], immersed=1 if sphere else 0, les=0, second_order=0, dpdz_set=1,
               ren=1000.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    coor = V.channel_grid(Ni + 1, Nj + 1, Nk + 1, W.LX, W.LY, W.LZ, beta, amp)
    o = Oracle(cfg, coor)
    if sphere:
        o.nvert_sphere(W.LX / 2, 1.0, 0.35, 0.3)
    rng = np.random.default_rng(5)
    ucat = rng.uniform(-1, 1, 3 * o.P)
    ucont = rng.uniform(-1, 1, 3 * o.P)
    return cfg, coor, o, ucat, ucont


def test_contra2cart_single_is_the_inverse_metric():
    rng = np.random.default_rng(3)
    for _ in range(20):
        M = rng.uniform(-1, 1, (3, 3)) + 2 * np.eye(3)
        uc = rng.uniform(-1, 1, 3)
        u = (C.c_double * 3)()
        a = [np.ascontiguousarray(M[r]) for r in range(3)]
        lib().vo_contra2cart_single(ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(uc), u)
        np.testing.assert_allclose(np.array(u[:]), np.linalg.solve(M, uc), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("sphere", [False, True])
def test_inlet_area_uniform_grid(sphere):
    cfg, coor, o, _, _ = open_case(sphere=sphere)
    Ni, Nj, Nk = cfg["mx"] - 2, cfg["my"] - 2, cfg["mz"] - 2
    dA = (W.LX / Ni) * (W.LY / Nj)
    ka, mean, inlet = o.inlet_area()
    nv = o.nvert.reshape(cfg["mz"], cfg["my"], cfg["mx"])
    for k in range(1, cfg["mz"] - 1):
        blocked = ((nv[k - 1, 1:-1, 1:-1] + nv[k, 1:-1, 1:-1]) >= 0.1).sum()
        assert abs(ka[k] - (Ni * Nj - blocked) * dA) < 1e-12
    assert ka[0] == ka[1]
    assert abs(mean - ka[1:-1].mean()) < 1e-12
    assert abs(inlet - ((nv[1, 1:-1, 1:-1] < 0.1).sum()) * dA) < 1e-12
    if sphere:
        assert ka.min() < W.LX * W.LY - 1e-9


def test_uniform_and_parabolic_inflow():
    cfg, coor, o, ucat, ucont = open_case()
    Ni, Nj = cfg["mx"] - 2, cfg["my"] - 2
    dx, dy = W.LX / Ni, W.LY / Nj
    ubcs = o.vec()
    f = o.inflow_flux(coor, ucont, ucat, ubcs, inletprofile=1)
    assert abs(f - W.LX * W.LY) < 1e-12
    mx, my = cfg["mx"], cfg["my"]
    uc = ucont.reshape(cfg["mz"], my, mx, 3)
    u = ucat.reshape(cfg["mz"], my, mx, 3)
    b = ubcs.reshape(cfg["mz"], my, mx, 3)
    assert np.abs(uc[0, 1:-1, 1:-1, 2] - dx * dy).max() < 1e-15
    assert np.all(uc[0, 1:-1, 1:-1, :2] == 0)
    avg = 0.5 * (u[0, 1:-1, 1:-1] + u[1, 1:-1, 1:-1])
    assert np.abs(avg - [0, 0, 1]).max() < 1e-14
    assert np.array_equal(b[0, 1:-1, 1:-1], u[0, 1:-1, 1:-1])
    f21 = o.inflow_flux(coor, ucont, ucat, ubcs, inletprofile=21)
    assert abs(f21 - W.LX * (2 + dy * dy / 4)) < 1e-12
    # flux scaling: inlet_flux / inlet_area
    f2 = o.inflow_flux(coor, ucont, ucat, ubcs, inletprofile=2, inlet_flux=3.0, inlet_area=W.LX * W.LY)
    assert abs(f2 - 3.0) < 1e-12


@pytest.mark.parametrize("sphere", [False, True])
def test_outlet_correction_balances_the_flux(sphere):
    cfg, coor, o, ucat, ucont = open_case(beta=1.5, amp=0.1, sphere=sphere)
    ubcs = o.vec()
    f_in = o.inflow_flux(coor, ucont, ucat, ubcs, inletprofile=21)
    before = o.outflow_flux(ucont)
    f_out, area = o.outlet_correction(f_in, ucat, ucont, ubcs)
    assert f_out == before and area > 0
    assert abs(o.outflow_flux(ucont) - f_in) < 1e-12 * max(1.0, abs(f_in))
    mz, my, mx = cfg["mz"], cfg["my"], cfg["mx"]
    u = ucat.reshape(mz, my, mx, 3)
    b = ubcs.reshape(mz, my, mx, 3)
    nv = o.nvert.reshape(mz, my, mx)
    fl = nv[mz - 2, 1:-1, 1:-1] < 0.1
    assert np.array_equal(b[mz - 1, 1:-1, 1:-1][fl], u[mz - 2, 1:-1, 1:-1][fl])
    assert np.all(b[mz - 1, 1:-1, 1:-1, 2][~fl] == 0)


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("profile", [1, 2, 10, 21, 0])
@pytest.mark.parametrize("sphere", [False, True])
def test_gpu_open_boundaries(profile, sphere):
    cfg, coor, o, ucat, ucont = open_case(N=(40, 24, 30), beta=1.5, amp=0.1, sphere=sphere)
    ctx = V.Ctx(cfg)
    dcoor = V.DeviceArray.from_host(coor)
    ctx.build_metrics(dcoor)
    ctx.set_nvert(V.DeviceArray.from_host(o.nvert))
    ka, mean, inlet = ctx.inlet_area()
    oka, omean, oinlet = o.inlet_area()
    # the device sums a plane in a fixed tree order, the oracle in the
    # reference's j, i loop: equal up to the rounding of the association
    assert rel_err(ka.host(), oka) <= 1e-14 and abs(mean - omean) <= 1e-14 * omean
    assert abs(inlet - oinlet) <= 1e-14 * oinlet
    plane = np.linspace(0.2, 1.3, cfg["mx"] * cfg["my"]) if profile == 0 else None
    kw = dict(inletprofile=profile, inlet_flux=2.5 if profile == 2 else -1.0, inlet_area=oinlet)
    d_uc, d_u, d_b = V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), ctx.vec()
    f = ctx.inflow_flux(dcoor, d_uc, d_u, d_b, uin_plane=None if plane is None else V.DeviceArray.from_host(plane), **kw)
    uc, u, b = ucont.copy(), ucat.copy(), o.vec()
    fo = o.inflow_flux(coor, uc, u, b, uin_plane=plane, **kw)
    tol = 1e-15 if profile == 10 else 0.0   # pow of the power law: the device's and the C library's
    for g, h in ((d_uc.host(), uc), (d_u.host(), u), (d_b.host(), b)):
        assert np.abs(g - h).max() <= tol * np.abs(h).max()
    # plane sums: fixed tree order on the device (see above)
    assert abs(f - fo) <= 1e-14 * abs(fo)
    assert abs(ctx.outflow_flux(d_uc) - o.outflow_flux(uc)) <= 1e-14 * np.abs(uc).max() * cfg["mx"] * cfg["my"]
    g_out = ctx.outlet_correction(fo, d_u, d_uc, d_b)
    o_out = o.outlet_correction(fo, u, uc, b)
    assert abs(g_out[1] - o_out[1]) <= 1e-14 * o_out[1]
    assert abs(g_out[0] - o_out[0]) <= 1e-14 * np.abs(uc).max() * cfg["mx"] * cfg["my"]
    for g, h in ((d_uc.host(), uc), (d_b.host(), b)):
        assert np.abs(g - h).max() <= 1e-14 * np.abs(h).max()


# ---------------------------------------------------------------------------
# the outlet rules inside the momentum routines
# ---------------------------------------------------------------------------
def outlet_case(bct5, sphere_at_outlet=True):
    N = (16, 12, 14)
    cfg = dict(mx=N[0] + 2, my=N[1] + 2, mz=N[2] + 2, i_periodic=1, j_periodic=0, k_periodic=0,
               bctype=[7, 7, 1, 1, 5, bct5], immersed=1, les=0, second_order=0, dpdz_set=1,
               ren=1000.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    coor = V.channel_grid(N[0] + 1, N[1] + 1, N[2] + 1, W.LX, W.LY, W.LZ, 1.5, 0.1)
    o = Oracle(cfg, coor)
    if sphere_at_outlet:
        o.nvert_sphere(W.LX / 2, 1.0, W.LZ - 0.2, 0.45)
    rng = np.random.default_rng(11)
    return cfg, coor, o, rng.uniform(-1, 1, 3 * o.P), rng.uniform(-1, 1, 3 * o.P)


@pytest.mark.parametrize("bct5", [4, 1])
def test_outlet_face_convects_with_the_flux_before_it(bct5):
    # momentum.c:1227-1229: at an outlet the last k-face of a fluid cell takes
    # its upwind direction from U^zeta of the face before it.  That direction
    # is used where the cell beyond is masked (the upwind branch), so with
    # ghost cells marked behind fluid last cells the rhs does not see U^zeta
    # of those last faces at an outlet, and does behind a wall
    cfg, coor, o, ucat, ucont = outlet_case(bct5, sphere_at_outlet=False)
    mz, my, mx = cfg["mz"], cfg["my"], cfg["mx"]
    nv = o.nvert.reshape(mz, my, mx)
    nv[mz - 1, 3:7, 2:9] = 2.0     # masked ghosts behind fluid last cells
    F = np.zeros((mz, my, mx), bool)
    F[mz - 2, 3:7, 2:9] = True
    assert np.all(nv[mz - 2][F[mz - 2]] == 0)
    rhs0 = o.vec()
    o.formfunction2(ucont.copy(), ucat.copy(), o.sca(), rhs0, 1.0)
    uc2 = ucont.copy().reshape(mz, my, mx, 3)
    uc2[..., 2][F] = -uc2[..., 2][F] + 0.37   # flips the upwind direction of most of them
    rhs1 = o.vec()
    o.formfunction2(uc2.ravel(), ucat.copy(), o.sca(), rhs1, 1.0)
    if bct5 == 4:
        assert np.array_equal(rhs0, rhs1)
    else:
        assert np.abs(rhs0 - rhs1).max() > 1e-6


def test_outlet_ghost_zero_behind_solid_cells():
    # rhs.c:646-663: the outlet ghost is zeroed behind a solid last cell and
    # otherwise left as it was
    cfg, coor, o, ucat, ucont = outlet_case(4)
    mz, my, mx = cfg["mz"], cfg["my"], cfg["mx"]
    before = ucat.copy().reshape(mz, my, mx, 3)
    u = ucat.copy()
    o.contra2cart(ucont.copy(), u)
    u = u.reshape(mz, my, mx, 3)
    nv = o.nvert.reshape(mz, my, mx)
    solid = nv[mz - 2, 1:-1, 1:-1] > 0.1
    assert solid.any() and (~solid).any()
    ghost = u[mz - 1, 1:-1, 1:-1]
    assert np.all(ghost[solid] == 0)
    assert np.array_equal(ghost[~solid], before[mz - 1, 1:-1, 1:-1][~solid])


@pytest.mark.gpu
@pytest.mark.parametrize("bct5", [4, 1])
def test_gpu_outlet_rules(bct5):
    # the outlet face rule and the outlet ghost rule against the oracle, on a
    # body at the outlet plus masked ghosts behind fluid last cells
    cfg, coor, o, ucat, ucont = outlet_case(bct5)
    mz, my, mx = cfg["mz"], cfg["my"], cfg["mx"]
    nv = o.nvert.reshape(mz, my, mx)
    nv[mz - 1, 3:7, 2:9] = np.where(nv[mz - 2, 3:7, 2:9] == 0, 2.0, nv[mz - 1, 3:7, 2:9])
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(o.nvert))
    d_uc, d_u, d_rhs = V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), ctx.vec()
    ctx.formfunction2(d_uc, d_u, ctx.sca(), d_rhs, 1.0)
    uc, u, rhs = ucont.copy(), ucat.copy(), o.vec()
    o.formfunction2(uc, u, o.sca(), rhs, 1.0)
    assert np.array_equal(d_rhs.host(), rhs)
    ctx.contra2cart(d_uc, d_u)
    o.contra2cart(uc, u)
    assert np.array_equal(d_u.host(), u)
