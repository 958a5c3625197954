"""Pins the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors and cannot be built
here (it needs PETSc 3.1, HYPRE 2.11 and MPI), so each oracle routine is
checked against closed forms, invariants and brute force chosen so that a
plausible restatement mistake (a dropped term, a wrong sign or index, a
transposed operand) fails at least one of them.
"""
import math
import os

import numpy as np
import pytest

from cases import oracle_case
from conftest import rel_err
from oracle_py import Oracle
from vfsgpu import workloads as W
import vfsgpu as V


def uniform_case(N=(12, 10, 8), immersed=0, les=0, all_periodic=False):
    Ni, Nj, Nk = N
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=1 if all_periodic else 0, k_periodic=1,
               bctype=[7, 7, 7, 7, 7, 7] if all_periodic else [7, 7, 1, 1, 7, 7], immersed=immersed, les=les,
               second_order=0, dpdz_set=1, ren=180.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    coor = V.channel_grid(Ni + 1, Nj + 1, Nk + 1, W.LX, W.LY, W.LZ, 0.0, 0.0)
    return cfg, coor, Oracle(cfg, coor)


def arr(o, a, comps=3):
    return a.reshape(o.mz, o.my, o.mx, comps) if comps > 1 else a.reshape(o.mz, o.my, o.mx)


# ---------------------------------------------------------------------------
# metrics (metrics.c)
# ---------------------------------------------------------------------------
def test_metrics_uniform_closed_form():
    cfg, coor, o = uniform_case()
    Ni, Nj, Nk = cfg["mx"] - 2, cfg["my"] - 2, cfg["mz"] - 2
    dx, dy, dz = W.LX / Ni, W.LY / Nj, W.LZ / Nk
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    csi, eta, zet = arr(o, o.g["csi"]), arr(o, o.g["eta"]), arr(o, o.g["zet"])
    aj = arr(o, o.g["aj"], 1)
    assert rel_err(csi[inner], np.broadcast_to([dy * dz, 0, 0], csi[inner].shape)) < 1e-14
    assert rel_err(eta[inner], np.broadcast_to([0, dz * dx, 0], eta[inner].shape)) < 1e-14
    assert rel_err(zet[inner], np.broadcast_to([0, 0, dx * dy], zet[inner].shape)) < 1e-14
    assert rel_err(aj[inner], np.full(aj[inner].shape, 1 / (dx * dy * dz))) < 1e-13
    # face metrics of a uniform grid equal the cell metrics; harmonic mean too
    for f, ref in (("icsi", [dy * dz, 0, 0]), ("jeta", [0, dz * dx, 0]), ("kzet", [0, 0, dx * dy])):
        a = arr(o, o.g[f])[inner]
        assert rel_err(a, np.broadcast_to(ref, a.shape)) < 1e-14, f
    for f in ("iaj", "jaj", "kaj"):
        a = arr(o, o.g[f], 1)[inner]
        assert rel_err(a, np.full(a.shape, 1 / (dx * dy * dz))) < 1e-13, f


def test_metrics_jacobian_is_inverse_cell_volume():
    """On the stretched, wavy C2 grid, 1/aj of every cell equals the volume of
    its hexahedron from the divergence theorem (independent formula)."""
    cfg, spec = W.make_config("C2", (10, 8, 6))
    coor = W.grid(spec)
    o = Oracle(cfg, coor)
    X = coor.reshape(o.mz, o.my, o.mx, 3)
    aj = arr(o, o.g["aj"], 1)
    for (k, j, i) in [(1, 1, 1), (3, 4, 5), (6, 8, 10), (2, 7, 3)]:
        c = np.array([[[X[k - 1 + a, j - 1 + b, i - 1 + q] for q in (0, 1)] for b in (0, 1)] for a in (0, 1)])
        # trilinear hexahedron volume via the 8-point Gauss rule of det(J)
        vol = 0.0
        g = 1 / math.sqrt(3)
        for s in (-g, g):
            for t in (-g, g):
                for u in (-g, g):
                    r, e, z = (s + 1) / 2, (t + 1) / 2, (u + 1) / 2
                    def d(f):
                        return f
                    Jm = np.zeros((3, 3))
                    # derivatives of the trilinear map w.r.t. (r, e, z)
                    Jm[:, 0] = ((1 - e) * (1 - z) * (c[0, 0, 1] - c[0, 0, 0]) + e * (1 - z) * (c[0, 1, 1] - c[0, 1, 0]) +
                                (1 - e) * z * (c[1, 0, 1] - c[1, 0, 0]) + e * z * (c[1, 1, 1] - c[1, 1, 0]))
                    Jm[:, 1] = ((1 - r) * (1 - z) * (c[0, 1, 0] - c[0, 0, 0]) + r * (1 - z) * (c[0, 1, 1] - c[0, 0, 1]) +
                                (1 - r) * z * (c[1, 1, 0] - c[1, 0, 0]) + r * z * (c[1, 1, 1] - c[1, 0, 1]))
                    Jm[:, 2] = ((1 - r) * (1 - e) * (c[1, 0, 0] - c[0, 0, 0]) + r * (1 - e) * (c[1, 0, 1] - c[0, 0, 1]) +
                                (1 - r) * e * (c[1, 1, 0] - c[0, 1, 0]) + r * e * (c[1, 1, 1] - c[0, 1, 1]))
                    vol += np.linalg.det(Jm) / 8
        # the reference's centred 8-node differences equal the Jacobian at the
        # cell centre, exact for the volume of a parallelepiped and second-order
        # accurate for a smooth hexahedron
        assert abs(1 / aj[k, j, i] - vol) / vol < 2e-3


def test_metrics_ghost_mirror_and_periodic():
    cfg, spec = W.make_config("C2", (10, 8, 6))
    o = Oracle(cfg, W.grid(spec))
    csi = arr(o, o.g["csi"])
    aj = arr(o, o.g["aj"], 1)
    # periodic in i and k: ghost 0 == cell m-2, ghost m-1 == cell 1
    assert np.array_equal(csi[3, 4, 0], csi[3, 4, o.mx - 2])
    assert np.array_equal(aj[0, 4, 3], aj[o.mz - 2, 4, 3])
    # mirrored in j (metrics.c:288-310)
    assert np.array_equal(aj[3, 0, 3], aj[3, 1, 3]) and np.array_equal(aj[3, o.my - 1, 3], aj[3, o.my - 2, 3])
    # face value is the average of the two cells, face Jacobian the harmonic mean
    icsi = arr(o, o.g["icsi"])
    iaj = arr(o, o.g["iaj"], 1)
    assert np.allclose(icsi[2, 3, 4], 0.5 * csi[2, 3, 4] + 0.5 * csi[2, 3, 5], rtol=0, atol=0)
    assert iaj[2, 3, 4] == 2.0 / (1.0 / aj[2, 3, 4] + 1.0 / aj[2, 3, 5])


# ---------------------------------------------------------------------------
# mask (ibm.c)
# ---------------------------------------------------------------------------
def test_nvert_sphere_brute_force():
    cfg, spec = W.make_config("C2", (24, 16, 12))
    coor = W.grid(spec)
    o = Oracle(cfg, coor)
    nv = arr(o, o.nvert_sphere(*W.SPHERE).copy(), 1)
    X = coor.reshape(o.mz, o.my, o.mx, 3)
    cent = 0.125 * (X[1:, 1:, 1:] + X[1:, :-1, 1:] + X[:-1, 1:, 1:] + X[:-1, :-1, 1:] +
                    X[1:, 1:, :-1] + X[1:, :-1, :-1] + X[:-1, 1:, :-1] + X[:-1, :-1, :-1])
    cx, cy, cz, r = W.SPHERE
    inside = np.zeros((o.mz, o.my, o.mx), bool)
    d = np.sqrt((cent[..., 0] - cx) ** 2 + (cent[..., 1] - cy) ** 2 + (cent[..., 2] - cz) ** 2) - r
    inside[1:-1, 1:-1, 1:-1] = d[:-1, :-1, :-1][: o.mz - 2, : o.my - 2, : o.mx - 2] < 0
    ref = np.zeros((o.mz, o.my, o.mx))
    ref[inside] = 3
    for k in range(1, o.mz - 1):
        for j in range(1, o.my - 1):
            for i in range(1, o.mx - 1):
                if not inside[k, j, i] and inside[k - 1:k + 2, j - 1:j + 2, i - 1:i + 2].any():
                    ref[k, j, i] = 1
    assert inside.sum() > 4
    assert np.array_equal(nv, ref)


# ---------------------------------------------------------------------------
# Contra2Cart (rhs.c)
# ---------------------------------------------------------------------------
def test_contra2cart_cramer_brute_force():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (12, 10, 8))
    u = ucat.copy()
    o.contra2cart(ucont.copy(), u)
    U = arr(o, u)
    Uc = arr(o, ucont)
    csi, eta, zet = arr(o, o.g["csi"]), arr(o, o.g["eta"]), arr(o, o.g["zet"])
    nv = arr(o, o.nvert, 1)
    for (k, j, i) in [(1, 1, 1), (4, 5, 6), (8, 10, 12), (3, 2, 7)]:
        if nv[k, j, i] >= 0.1:
            continue
        M = np.array([csi[k, j, i], eta[k, j, i], zet[k, j, i]])
        q = np.array([0.5 * (Uc[k, j, i - 1, 0] + Uc[k, j, i, 0]), 0.5 * (Uc[k, j - 1, i, 1] + Uc[k, j, i, 1]),
                      0.5 * (Uc[k - 1, j, i, 2] + Uc[k, j, i, 2])])
        assert rel_err(U[k, j, i], np.linalg.solve(M, q)) < 1e-12


def test_contra2cart_noslip_ghosts_and_linearity():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C1", (10, 8, 6))
    u1 = ucat.copy(); o.contra2cart(ucont.copy(), u1)
    U = arr(o, u1)
    # no-slip mirror: ghost = -first interior cell (rhs.c:554-567)
    assert np.array_equal(U[3, 0, 4], -U[3, 1, 4]) and np.array_equal(U[3, o.my - 1, 4], -U[3, o.my - 2, 4])
    # periodic ghosts
    assert np.array_equal(U[3, 4, 0], U[3, 4, o.mx - 2])
    # linear in ucont
    u2 = ucat.copy(); o.contra2cart(2.5 * ucont, u2)
    assert rel_err(u2, 2.5 * u1) < 1e-14


# ---------------------------------------------------------------------------
# wall-function walls (bctype -1): the Cabot wall model of wallfunction.c
# ---------------------------------------------------------------------------
def _wf_integrand(yp):
    """1 / (1 + nu_t/nu), nu_t/nu = kappa y+ (1 - exp(-y+/19))^2, kappa 0.41."""
    return 1.0 / (1.0 + 0.41 * yp * (1.0 - math.exp(-yp / 19.0)) ** 2)


def test_wall_table_is_the_mixing_length_integral():
    """table[n] = integral of 1/(1 + nu_t/nu) over y+ in [0, 2n], against
    adaptive quadrature; in the log region the integral has the closed form
    ln((1 + kappa b) / (1 + kappa a)) / kappa."""
    from scipy.integrate import quad
    import oracle_py as OP
    tab = OP.wf_table()
    assert tab[0] == 0.0 and tab.size == 500001
    for n in (1, 3, 10, 50, 500):
        want = quad(_wf_integrand, 0.0, 2.0 * n, epsabs=0, epsrel=1e-13, limit=400)[0]
        assert abs(tab[n] - want) <= 1e-10 * want, n
    for a, b in ((2000, 20000), (20000, 1000000)):
        want = math.log((1 + 0.41 * b) / (1 + 0.41 * a)) / 0.41
        assert abs((tab[b // 2] - tab[a // 2]) - want) <= 1e-10 * want


def test_wall_model_utau_and_profile():
    """find_utau inverts u = utau^2 integrate_F(utau, y): in the viscous
    sublayer utau = sqrt(u nu / y) up to the table's linear first panel, and
    on table nodes integrate_F is exactly table / utau."""
    import oracle_py as OP
    tab = OP.wf_table()
    nu, y = 1e-3, 1e-4
    u = 0.25 * y / nu                                   # utau = 0.5 gives y+ = 0.05
    ut = OP.wf_find_utau(nu, u, y)
    slope = tab[1] / 2.0                                # linear interpolation on [0, 2]
    assert abs(ut - math.sqrt(u * nu / (y * slope))) <= 1e-9 * ut
    assert abs(ut - 0.5) <= 2.5e-3
    # a node: y+ = 2000 exactly
    utau, yy = 0.05, 2000 * nu / 0.05
    assert OP.wf_integrate_F(nu, utau, yy) == tab[1000] / utau
    # Newton recovers utau from the velocity it implies (log region)
    U = utau * utau * OP.wf_integrate_F(nu, utau, yy)
    assert abs(OP.wf_find_utau(nu, U, yy) - utau) <= 1e-9


def test_wall_function_splits_tangent_and_normal():
    """Ub: tangential part along Uc's with |Ub_t| = u_Cabot(sb, utau), utau
    from |Uc_t| at sc; normal part sb/sc times Uc's."""
    import oracle_py as OP
    rng = np.random.default_rng(3)
    n = rng.standard_normal(3); n /= np.linalg.norm(n)
    Uc = rng.standard_normal(3)
    nu, sb, sc = 1 / 5000.0, 0.004, 0.012
    Ub, us = OP.wall_function(nu, sc, sb, Uc, n)
    un = Uc @ n
    ut_c = Uc - un * n
    assert abs(us - OP.wf_find_utau(nu, np.linalg.norm(ut_c), sc)) <= 1e-15 * us
    ub_n = Ub @ n
    ub_t = Ub - ub_n * n
    assert abs(ub_n - sb / sc * un) <= 1e-13 * abs(un)
    assert abs(np.linalg.norm(ub_t) - us * us * OP.wf_integrate_F(nu, us, sb)) <= 1e-13
    assert np.linalg.norm(np.cross(ub_t, ut_c)) <= 1e-13 * np.linalg.norm(ub_t) * np.linalg.norm(ut_c)


def test_wall_cells_on_a_uniform_channel():
    """Uniform Cartesian channel with a wall-function bottom: the first
    cells are IB nodes, their velocity is the wall model of the second cell's
    with sb = dy/2, sc = 3dy/2 and the normal +y, the top wall keeps the
    no-slip mirror, and IB_BC zeroes the bottom wall flux."""
    import oracle_py as OP
    cfg, coor, o0 = uniform_case()
    cfg["bctype"] = [7, 7, -1, 1, 7, 7]
    cfg["ren"] = 5000.0
    o = Oracle(cfg, coor)
    o.wall_mask()
    nv = arr(o, o.nvert, 1)
    assert np.all(nv[1:-1, 1, 1:-1] == 1) and nv[1:-1, 2:-1, 1:-1].sum() == 0 and nv[:, 0].sum() == 0
    ucat = W.random_vector(cfg, 0.5, 71)
    ucont = W.random_vector(cfg, 0.5, 72)
    us = o.sca()
    o.contra2cart_ustar(ucont.copy(), ucat, us)
    U, S = arr(o, ucat), arr(o, us, 1)
    dy = W.LY / (o.my - 2)
    for (k, i) in [(1, 1), (3, 5), (o.mz - 2, o.mx - 2)]:
        Ub, ust = OP.wall_function(1 / 5000.0, 1.5 * dy, 0.5 * dy, U[k, 2, i], np.array([0.0, 1.0, 0.0]))
        assert rel_err(U[k, 1, i], Ub) < 1e-12 and abs(S[k, 1, i] - ust) <= 1e-12 * ust
    assert np.array_equal(U[3, o.my - 1, 4], -U[3, o.my - 2, 4])
    assert S[:, 2:].sum() == 0
    uc = W.random_vector(cfg, 0.5, 73)
    o.ib_bc(uc, ucat)
    assert not arr(o, uc)[1:-1, 0, 1:-1, 1].any()


# ---------------------------------------------------------------------------
# Formfunction_2 (momentum.c)
# ---------------------------------------------------------------------------
def test_ff2_poiseuille_closed_form():
    """u = 1-(y-1)^2: convection cancels and the viscous flux difference is
    exact for quadratics, so interior rhs.x = nu * u'' * dx*dy*dz * csi.x*iaj
    = -2 nu dy dz; rhs.y = rhs.z = 0."""
    Ni, Nj, Nk = 16, 16, 16
    cfg, coor, o = uniform_case((Ni, Nj, Nk))
    dx, dy, dz = W.LX / Ni, W.LY / Nj, W.LZ / Nk
    ucat, ucont = V.poiseuille(cfg, W.LY, 0.0, 1, o.g["cent"], o.nvert, o.g["icsi"], o.g["jeta"], o.g["kzet"])
    u = ucat.copy(); uc = ucont.copy()
    o.contra2cart(uc, u)
    rhs = o.vec()
    o.formfunction2(uc, u, o.sca(), rhs, 1.0)
    R = arr(o, rhs)
    nu = 1.0 / cfg["ren"]
    interior = R[1:-1, 3:-3, 1:-1]
    assert rel_err(interior[..., 0], np.full(interior[..., 0].shape, -2 * nu * dy * dz)) < 1e-10
    assert np.abs(R[..., 1]).max() < 1e-15 and np.abs(R[..., 2]).max() < 1e-15


def test_ff2_inviscid_poiseuille_has_zero_rhs():
    """-inv drops the viscous term (momentum.c:1621): the same Poiseuille field
    whose viscous rhs.x is -2 nu dy dz above gives zero, since its WENO3 face
    values are i-invariant and the transverse fluxes vanish."""
    Ni, Nj, Nk = 16, 16, 16
    cfg, coor, o = uniform_case((Ni, Nj, Nk))
    cfg = dict(cfg, inviscid=1)
    o = Oracle(cfg, coor)
    ucat, ucont = V.poiseuille(cfg, W.LY, 0.0, 1, o.g["cent"], o.nvert, o.g["icsi"], o.g["jeta"], o.g["kzet"])
    u = ucat.copy(); uc = ucont.copy()
    o.contra2cart(uc, u)
    rhs = o.vec()
    o.formfunction2(uc, u, o.sca(), rhs, 1.0)
    nu, dy, dz = 1.0 / cfg["ren"], W.LY / Nj, W.LZ / Nk
    assert np.abs(rhs).max() < 1e-12 * 2 * nu * dy * dz   # against the viscous value: roundoff only


def test_ff2_uniform_flow_has_zero_rhs():
    """Triply periodic uniform grid, constant velocity: every face flux is the
    same, so divergence, 4th-order correction and viscous terms vanish."""
    cfg, coor, o = uniform_case((8, 8, 8), all_periodic=True)
    ucat = np.zeros(3 * o.P)
    U = np.array([1.0, -0.3, 0.7])
    ucat.reshape(-1, 3)[:] = U
    ucont = np.zeros(3 * o.P)
    for f, q in (("icsi", 0), ("jeta", 1), ("kzet", 2)):
        A = o.g[f].reshape(-1, 3)
        ucont.reshape(-1, 3)[:, q] = A @ U
    u = ucat.copy()
    o.contra2cart(ucont, u)
    rhs = o.vec()
    o.formfunction2(ucont, u, o.sca(), rhs, 1.0)
    assert np.abs(rhs).max() < 1e-14


def test_ff2_accumulates_and_scales():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (12, 10, 8))
    u = ucat.copy(); uc = ucont.copy()
    o.contra2cart(uc, u)
    nut = o.sca(); o.eddy_viscosity(u, W.CS_LES, nut)
    r1 = o.vec(); o.formfunction2(uc, u, nut, r1, 1.0)
    r2 = np.ones(3 * o.P); o.formfunction2(uc, u, nut, r2, 0.5)
    R1, R2 = arr(o, r1), arr(o, r2)
    m = R1 != 0
    # rhs += scale*F on unmasked interior faces
    assert rel_err(R2[m], 1.0 + 0.5 * R1[m]) < 1e-14


# ---------------------------------------------------------------------------
# WENO3 (level.c:1734-1771, -weno 1)
# ---------------------------------------------------------------------------
def weno3(f0, f1, f2, f3, ws):
    from oracle_py import lib as olib
    import ctypes as C
    fn = olib().vo_weno3
    fn.restype = C.c_double
    fn.argtypes = [C.c_double] * 5
    return fn(f0, f1, f2, f3, ws)


def test_weno3_known_answers():
    rng = np.random.default_rng(5)
    for _ in range(50):
        a, b = rng.uniform(-2, 2, 2)
        f = [a + b * t for t in (-1.0, 0.0, 1.0, 2.0)]
        for ws in (1.0, -1.0):
            # linear data: both 2-point candidates equal the face value a + b/2
            assert abs(weno3(*f, ws) - (a + 0.5 * b)) <= 1e-14 * (abs(a) + abs(b) + 1)
        c = rng.uniform(-1, 1)
        assert abs(weno3(c, c, c, c, 1.0) - c) <= 1e-15 * (1 + abs(c))
    # a step: the stencil that does not cross it wins (upwind side by wave speed)
    assert abs(weno3(0.0, 0.0, 1.0, 1.0, 1.0) - 0.0) < 1e-10
    assert abs(weno3(0.0, 0.0, 1.0, 1.0, -1.0) - 1.0) < 1e-10
    # wave speed > 0 uses (f0, f1, f2); < 0 mirrors to (f3, f2, f1)
    assert weno3(0.3, -0.2, 0.9, 5.0, -1.0) == weno3(123.0, -0.2, 0.9, 5.0, -1.0)
    assert weno3(0.3, -0.2, 0.9, 5.0, 1.0) == weno3(0.3, -0.2, 0.9, -42.0, 1.0)


def test_weno3_square_is_pow():
    """The reference squares with pow(x, 2.) and is built with -O3
    (makefile:19), where GCC folds pow(x, 2.) into x * x.  Compile that
    expression the same way and check the restatement's x * x against it."""
    import ctypes as C
    import subprocess
    import tempfile
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "p.c"), "w") as f:
        f.write("#include <math.h>\ndouble sqpow(double x) { return pow(x, 2.); }\n")
    subprocess.run(["gcc", "-O3", "-shared", "-fPIC", "-o", os.path.join(d, "p.so"), os.path.join(d, "p.c"), "-lm"],
                   check=True)
    fn = C.CDLL(os.path.join(d, "p.so")).sqpow
    fn.restype = C.c_double
    fn.argtypes = [C.c_double]
    rng = np.random.default_rng(6)
    for x in np.concatenate([rng.uniform(-3, 3, 2000), rng.uniform(-1e-3, 1e-3, 2000) + 1e-6]):
        assert fn(float(x)) == float(x) * float(x)


def test_ff2_weno_equals_central_on_linear_fields():
    """On a uniform grid a field linear in x, y, z is reproduced exactly by
    both 2-point WENO candidates and by the 4th-order central formula, so the
    two schemes agree away from the boundary faces (where they differ)."""
    N = (14, 12, 12)
    Ni, Nj, Nk = N
    base = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=0, j_periodic=0, k_periodic=0,
                bctype=[1, 1, 1, 1, 1, 1], immersed=0, les=0, second_order=0, dpdz_set=1, ren=100.0,
                dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    coor = V.channel_grid(Ni + 1, Nj + 1, Nk + 1, 1.4, 1.2, 1.2, 0.0, 0.0)
    out = {}
    for weno in (0, 1):
        cfg = dict(base, weno=weno)
        o = Oracle(cfg, coor)
        # linear in the cell index = linear in space on a uniform grid, ghosts included
        kk, jj, ii = np.meshgrid(np.arange(o.mz), np.arange(o.my), np.arange(o.mx), indexing="ij")
        U = np.empty((o.mz, o.my, o.mx, 3))
        A = np.array([[0.3, 0.05, -0.02, 0.01], [-0.4, 0.02, 0.03, -0.05], [0.2, -0.01, 0.04, 0.06]])
        for q in range(3):
            U[..., q] = A[q, 0] + A[q, 1] * ii + A[q, 2] * jj + A[q, 3] * kk
        ucont = W.random_vector(cfg, 1.0, 77)
        rhs = o.vec()
        o.formfunction2(ucont.copy(), U.reshape(-1).copy(), o.sca(), rhs, 1.0)
        out[weno] = arr(o, rhs)
    inner = (slice(4, -4), slice(4, -4), slice(4, -4))
    a, b = out[1][inner], out[0][inner]
    assert rel_err(a, b) < 1e-12
    # and the WENO branch is really taken: near the walls the schemes differ
    assert np.abs(out[1] - out[0]).max() > 1e-8 * np.abs(out[0]).max()


# ---------------------------------------------------------------------------
# LES (les.c)
# ---------------------------------------------------------------------------
def test_les_linear_shear_closed_form():
    Ni, Nj, Nk = 10, 12, 8
    cfg, coor, o = uniform_case((Ni, Nj, Nk), les=1)
    dx, dy, dz = W.LX / Ni, W.LY / Nj, W.LZ / Nk
    gamma = 0.7
    ucat = np.zeros(3 * o.P)
    cent = o.g["cent"].reshape(-1, 3)
    ucat.reshape(-1, 3)[:, 0] = gamma * cent[:, 1]
    U = arr(o, ucat)
    U[:, 0] = -U[:, 1]; U[:, -1] = -U[:, -2]
    U[:, :, 0] = U[:, :, -2]; U[:, :, -1] = U[:, :, 1]   # periodic ghosts
    U[0] = U[-2]; U[-1] = U[1]
    nut = o.sca()
    o.eddy_viscosity(ucat, 0.01, nut)
    N = arr(o, nut, 1)[1:-1, 2:-2, 1:-1]
    ref = 0.01 * (dx * dy * dz) ** (2 / 3) * abs(gamma)
    assert rel_err(N, np.full(N.shape, ref)) < 1e-12


# ---------------------------------------------------------------------------
# Pressure gradient (momentum.c:197-440)
# ---------------------------------------------------------------------------
def test_pressure_gradient_linear_pressure():
    """p = a*y on a uniform grid: dP.y = g22 * a*dy * jaj = a * dz*dx, dP.x = dP.z = 0."""
    Ni, Nj, Nk = 10, 12, 8
    cfg, coor, o = uniform_case((Ni, Nj, Nk))
    dx, dy, dz = W.LX / Ni, W.LY / Nj, W.LZ / Nk
    a = 0.37
    p = a * o.g["cent"].reshape(-1, 3)[:, 1].copy()
    dp = o.vec()
    o.pressure_gradient(p, dp)
    D = arr(o, dp)
    inner = D[1:-1, 1:-2, 1:-1]
    assert rel_err(inner[..., 1], np.full(inner[..., 1].shape, a * dz * dx)) < 1e-12
    assert np.abs(D[..., 0]).max() < 1e-14 and np.abs(D[..., 2]).max() < 1e-14
    assert np.all(D[:, o.my - 2, :, 1] == 0)


# ---------------------------------------------------------------------------
# Poisson operator (poisson.c:971-1682)
# ---------------------------------------------------------------------------
def test_poisson_uniform_reduces_to_7_point():
    Ni, Nj, Nk = 8, 6, 5
    cfg, coor, o = uniform_case((Ni, Nj, Nk))
    dx, dy, dz = W.LX / Ni, W.LY / Nj, W.LZ / Nk
    coef = o.poisson_stencil().reshape(o.mz, o.my, o.mx, 19)
    c = coef[2, 3, 4]
    EP, WP, NP, SP, TP, BP = c[1], c[2], c[3], c[4], c[5], c[6]
    assert rel_err([EP, WP], [dy * dz / dx] * 2) < 1e-13
    assert rel_err([NP, SP], [dz * dx / dy] * 2) < 1e-13
    assert rel_err([TP, BP], [dx * dy / dz] * 2) < 1e-13
    assert np.abs(c[7:]).max() < 1e-12
    assert abs(c[0] + c[1:7].sum()) < 1e-12 * abs(c[0])
    # wall row: the wall neighbour is missing (Neumann)
    w = coef[2, 1, 4]
    assert w[4] == 0 and abs(w[0] + w[1:7].sum()) < 1e-12 * abs(w[0])


def _csr_dense(o):
    rp, col, val = o.poisson_lhs()
    n = o.n
    A = np.zeros((n, n))
    for r in range(n):
        for q in range(rp[r], rp[r + 1]):
            A[r, col[q]] += val[q]
    return A


def test_poisson_conservation_and_nullspace_curvilinear():
    """Flux form on the curvilinear grid with the sphere: A 1 = 0 (constants
    are the null space) and 1^T A = 0 (face fluxes cancel pairwise)."""
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (12, 10, 8))
    o.setup_gid()
    A = _csr_dense(o)
    scale = np.abs(A).max()
    assert np.abs(A.sum(axis=1)).max() < 1e-13 * scale
    assert np.abs(A.sum(axis=0)).max() < 1e-13 * scale


def test_poisson_uniform_operator_symmetric():
    cfg, coor, o = uniform_case((8, 6, 5))
    o.setup_gid()
    A = _csr_dense(o)
    assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()


def test_poisson_linear_field_on_sheared_grid():
    """On an affine (sheared) grid the metrics are constant and exact, so the
    19-point operator applied to any linear pressure vanishes on rows whose
    stencil stays clear of walls and periodic seams: this exercises every
    cross-derivative coefficient and its sign."""
    Ni, Nj, Nk = 10, 10, 10
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=0, k_periodic=1,
               bctype=[7, 7, 1, 1, 7, 7], immersed=0, les=0, second_order=0, dpdz_set=1, ren=100.0, dt=1e-3,
               st=1.0, mean_pressure_gradient=0.0)
    coor = np.zeros(3 * (Ni + 2) * (Nj + 2) * (Nk + 2))
    X = coor.reshape(Nk + 2, Nj + 2, Ni + 2, 3)
    kk, jj, ii = np.meshgrid(np.arange(Nk + 1), np.arange(Nj + 1), np.arange(Ni + 1), indexing="ij")
    xi, et, ze = ii * 0.1, jj * 0.13, kk * 0.09
    X[:Nk + 1, :Nj + 1, :Ni + 1, 0] = xi + 0.31 * et + 0.05 * ze
    X[:Nk + 1, :Nj + 1, :Ni + 1, 1] = et + 0.17 * ze + 0.08 * xi
    X[:Nk + 1, :Nj + 1, :Ni + 1, 2] = ze + 0.23 * xi - 0.11 * et
    o = Oracle(cfg, coor)
    cent = o.g["cent"].reshape(-1, 3)
    phi = 0.7 * cent[:, 0] - 1.3 * cent[:, 1] + 0.4 * cent[:, 2]
    o.setup_gid()
    o.poisson_lhs()
    y = o.sca()
    o.csr_apply(phi.copy(), y)
    Y = arr(o, y, 1)
    inner = Y[3:-3, 3:-3, 3:-3]
    coef = o.poisson_stencil().reshape(o.mz, o.my, o.mx, 19)
    scale = np.abs(coef[3:-3, 3:-3, 3:-3]).max() * np.abs(phi).max()
    assert np.abs(coef[5, 5, 5, 7:]).max() > 1e-6   # cross terms present
    assert np.abs(inner).max() < 1e-12 * scale


# ---------------------------------------------------------------------------
# projection chain (poisson_hypre.c, poisson.c, main.c)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["C1", "C2"])
def test_projection_removes_divergence(name):
    cfg, spec, coor, o, ucat, ucont = oracle_case(name, (12, 10, 8))
    uc = ucont.copy()
    if spec["immersed"]:
        o.volume_flux(uc)
    b = o.sca()
    o.poisson_rhs(uc, b)
    d0, _ = o.divergence(uc)
    phi = o.sca()
    it, rel = o.poisson_solve(b, phi, rtol=1e-13)
    assert rel < 1e-12
    o.projection(phi, uc)
    d1, _ = o.divergence(uc)
    assert d1 < 1e-9 * d0


def test_volume_flux_balances_ib_faces():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (24, 16, 12))
    uc = ucont.copy()
    flux, area = o.volume_flux(uc)
    assert area > 0
    assert abs(flux) < 1e-13 * area


def test_volume_flux_without_immersed_zeroes_mask_faces():
    """poisson.c:3306-3500 (immersed == 0): every face between a fluid cell
    (nvert < 0.1) and a marked one (> 0.1) ends with zero flux; the other
    faces are untouched; the returned flux is 0 and the area the sum of the
    cell-metric norms on those faces (brute force)."""
    cfg, coor, o = uniform_case((10, 8, 8))
    nv = np.zeros((o.mz, o.my, o.mx))
    nv[3:6, 2:5, 4:7] = 1.0
    nv[4, 3, 5] = 3.0
    o.nvert[:] = nv.ravel()
    uc = W.random_vector(cfg, 1.0, 91)
    u = uc.copy()
    f, a = o.volume_flux(u)
    U, U0 = arr(o, u), arr(o, uc)
    met = [arr(o, o.g[n]) for n in ("csi", "eta", "zet")]
    area = 0.0
    changed = np.zeros(U.shape, bool)
    for k in range(1, o.mz - 1):
        for j in range(1, o.my - 1):
            for i in range(1, o.mx - 1):
                nbs = [nv[k, j, i + 1], nv[k, j + 1, i], nv[k + 1, j, i]]
                for q in range(3):
                    if (nv[k, j, i] < 0.1) != (nbs[q] < 0.1):
                        changed[k, j, i, q] = True
                        area += np.linalg.norm(met[q][k, j, i])
    assert np.all(U[changed] == 0)
    assert np.array_equal(U[~changed], U0[~changed])
    assert abs(f) == 0.0 and rel_err(a, area) < 1e-13
    # nvert == 0 everywhere: nothing to do
    o.nvert[:] = 0
    u = uc.copy()
    f, a = o.volume_flux(u)
    assert np.array_equal(u, uc) and f == 0 and a == 0


def test_poisson_rhs_has_zero_sum():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (12, 10, 8))
    b = o.sca()
    o.poisson_rhs(ucont.copy(), b)
    assert abs(b.sum()) < 1e-10 * np.abs(b).sum()


def test_update_pressure_rules():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (24, 16, 12))
    p = np.full(o.P, 2.0)
    phi = np.arange(o.P, dtype=float) * 1e-3
    p0 = p.copy()
    o.update_pressure(p, phi)
    nv = o.nvert
    P = arr(o, p, 1)[1:-1, 1:-1, 1:-1]
    N = arr(o, nv, 1)[1:-1, 1:-1, 1:-1]
    PH = arr(o, phi, 1)[1:-1, 1:-1, 1:-1]
    assert np.all(P[N > 1.1] == 0)
    assert np.allclose(P[N < 0.1], 2.0 + PH[N < 0.1], rtol=0, atol=0)
    assert np.all(P[(N > 0.1) & (N < 1.1)] == 2.0)
