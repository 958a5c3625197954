"""Pins of the oracle's momentum and pressure routines that the uniform-grid
closed forms of test_oracle.py leave open: the curvilinear cross-metric
terms, the branches next to an immersed body and at walls, the eddy
viscosity selection at IB faces, Calc_Minimum_dt, the mean pressure
gradient, IB_BC's face rule, and the reference's own known answer, the 2-D
Taylor-Green vortex (bcs.c:3245-3300, rhs2.c:1358-1428).

Each check is a function pin_*() that raises AssertionError.  The last test
compiles copies of the oracle with one planted slip each (a transposed
index, a dropped term, a wrong coefficient, a stencil that reads the wrong
side of a body) and shows that every slip fails at least one pin.

Geometry used below.  On an affine grid X = M (i, j, k) the metrics are
exact and constant: the area vectors of the i, j, k faces are
S_i = e_eta x e_zeta, S_j = e_zeta x e_xi, S_k = e_xi x e_eta (e_* the
columns of M) and the cell volume is det M, with sum_m S_m (x) e_m = V I.
With that identity,
- a linear pressure p = g.x gives the face gradient S.g on every face;
- a linear velocity u = G x + u0 has the viscous face flux
  nu (G + G^T) S on every face, so the viscous part of the right-hand side
  vanishes, and the convective right-hand side at face f is
  -S_f . (G u + u tr G)(x_f), since the face fluxes are quadratic along the
  lattice and both the 4th-order and the 2nd-order differences are exact
  for them;
- a quadratic velocity has the viscous right-hand side
  nu S_f . (lap u + grad div u), constant in space.
"""
import contextlib
import ctypes as C
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_py as OP
from cases import oracle_case
from conftest import rel_err
from oracle_py import Oracle
from vfsgpu import workloads as W
import vfsgpu as V

M_AFF = np.array([[0.100, 0.031, -0.012],
                  [0.014, 0.130, 0.021],
                  [0.019, -0.017, 0.090]])   # columns: e_xi, e_eta, e_zeta


# ---------------------------------------------------------------------------
# affine grids
# ---------------------------------------------------------------------------
class Affine:
    def __init__(self, N=(12, 11, 10), M=M_AFF, per=(0, 0, 0), ren=100.0, immersed=0, les=0, mpg=0.0, dpdz_set=1):
        Ni, Nj, Nk = N
        self.M = np.asarray(M, float)
        self.mx, self.my, self.mz = Ni + 2, Nj + 2, Nk + 2
        kk, jj, ii = np.meshgrid(np.arange(self.mz), np.arange(self.my), np.arange(self.mx), indexing="ij")
        self.idx = np.stack([ii, jj, kk], -1).astype(float)
        coor = self.idx @ self.M.T
        bct = [7 if per[d // 2] else 1 for d in range(6)]
        self.cfg = dict(mx=self.mx, my=self.my, mz=self.mz, i_periodic=per[0], j_periodic=per[1], k_periodic=per[2],
                        bctype=bct, immersed=immersed, les=les, second_order=0, dpdz_set=dpdz_set, weno=0, inviscid=0,
                        ren=ren, dt=1e-3, st=1.0, mean_pressure_gradient=mpg)
        self.o = Oracle(self.cfg, np.ascontiguousarray(coor.ravel()))
        e = [self.M[:, m] for m in range(3)]
        self.S = [np.cross(e[1], e[2]), np.cross(e[2], e[0]), np.cross(e[0], e[1])]
        self.vol = float(np.linalg.det(self.M))
        # cell centres and face centres of cell (k, j, i)
        self.xc = (self.idx - 0.5) @ self.M.T
        self.xf = []
        for d in range(3):
            sh = np.full(3, -0.5)
            sh[d] = 0.0
            self.xf.append((self.idx + sh) @ self.M.T)

    def face_flux(self, ufun):
        """ucont = u(x_f) . S_f on every face (the exact contravariant flux)."""
        uc = np.zeros((self.mz, self.my, self.mx, 3))
        for d in range(3):
            uc[..., d] = ufun(self.xf[d]) @ self.S[d]
        return uc

    def cells(self, ufun):
        return ufun(self.xc)

    def planar_body(self, axis, high=False, nb=3):
        """Solid (3) for the ghost and the first nb cell layers along `axis`
        (the last ones with high), an IB layer (1) next to them, fluid (0)
        elsewhere."""
        nv = np.zeros((self.mz, self.my, self.mx))
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(0, nb + 1)
        nv[tuple(sl)] = 3.0
        sl[2 - axis] = nb + 1
        nv[tuple(sl)] = 1.0
        return np.flip(nv, axis=2 - axis).copy() if high else nv

    def bodies(self):
        return [self.planar_body(a, h) for a in range(3) for h in (False, True)]


def _sphere_mask(aff, r=0.2):
    """nvert of a sphere by the SDF rule (solid 3, its 3x3x3 shell 1); on a
    16^3 grid it stays four cells clear of the walls, as a body in the
    reference does (a stencil switched to one side of an IB cell next to a
    wall would read the wall's ghost)."""
    c = aff.xc[aff.mz // 2, aff.my // 2, aff.mx // 2]
    d = np.linalg.norm(aff.xc - c, axis=-1) - r
    solid = np.zeros(d.shape, bool)
    solid[1:-1, 1:-1, 1:-1] = d[1:-1, 1:-1, 1:-1] < 0
    near = np.zeros_like(solid)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for e in (-1, 0, 1):
                near |= np.roll(np.roll(np.roll(solid, a, 0), b, 1), e, 2)
    nv = np.where(solid, 3.0, np.where(near, 1.0, 0.0))
    nv[0], nv[-1], nv[:, 0], nv[:, -1], nv[:, :, 0], nv[:, :, -1] = 0, 0, 0, 0, 0, 0
    return nv


def _unmasked(aff, nv, d):
    """Faces d of cell (k,j,i) that Formfunction_2 and Pressure_Gradient
    leave unmasked: interior cell, nvert sum <= 0.1, not the last cell of a
    non-periodic direction (momentum.c:1795-1803, 415-425)."""
    m = np.zeros(nv.shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    nb = np.roll(nv, -1, axis=2 - d)
    m &= (nv + nb) <= 0.1
    n = [aff.mx, aff.my, aff.mz][d]
    if not aff.cfg[("i_periodic", "j_periodic", "k_periodic")[d]]:
        sl = [slice(None)] * 3
        sl[2 - d] = n - 2
        m[tuple(sl)] = False
    return m


# ---------------------------------------------------------------------------
# pins
# ---------------------------------------------------------------------------
def pin_pressure_gradient_affine_with_body():
    """Linear p on the sheared grid, p garbage in IB, solid and ghost cells:
    dP = S.g on every unmasked face, 0 on the masked ones (the one-sided
    branches next to IB cells and walls, the cross-metric terms)."""
    rng = np.random.default_rng(11)
    g = np.array([0.7, -1.3, 0.45])
    big, aff = Affine((16, 16, 16), immersed=1), Affine(immersed=1)
    for a, nv in [(big, _sphere_mask(big))] + [(aff, b) for b in aff.bodies()]:
        o = a.o
        o.nvert[:] = nv.ravel()
        p = a.xc @ g + 0.2
        bad = nv > 0.5
        bad[0], bad[-1], bad[:, 0], bad[:, -1], bad[:, :, 0], bad[:, :, -1] = (True,) * 6
        p[bad] = rng.uniform(-50, 50, bad.sum())
        dp = o.vec()
        o.pressure_gradient(np.ascontiguousarray(p.ravel()), dp)
        D = dp.reshape(a.mz, a.my, a.mx, 3)
        for d in range(3):
            m = _unmasked(a, nv, d)
            ref = a.S[d] @ g
            assert m.sum() > 50
            assert np.abs(D[..., d][m] - ref).max() <= 1e-12 * abs(ref), d
            assert not D[..., d][~m].any(), d


def pin_mean_pressure_gradient():
    """k-periodic uniform grid, p = g.x: dP.z = (g_z + mean_pressure_gradient)
    dx dy (momentum.c:382-386); dP.x and dP.y are unchanged by it."""
    Ni, Nj, Nk = 8, 9, 10
    dx, dy, dz = 0.3, 0.2, 0.25
    M = np.diag([dx, dy, dz])
    g = np.array([0.2, -0.4, 0.3])
    out = {}
    for mpg in (0.0, -1.7):
        aff = Affine((Ni, Nj, Nk), M, per=(1, 0, 1), mpg=mpg)
        o = aff.o
        p = aff.xc @ g
        # periodic in i and k: a linear p is not, so only faces away from the
        # seams are compared
        dp = o.vec()
        o.pressure_gradient(np.ascontiguousarray(p.ravel()), dp)
        out[mpg] = dp.reshape(aff.mz, aff.my, aff.mx, 3)
    inner = (slice(2, -3), slice(1, -2), slice(2, -3))
    a, b = out[-1.7][inner], out[0.0][inner]
    assert np.abs(b[..., 2] - g[2] * dx * dy).max() <= 1e-12 * abs(g[2]) * dx * dy
    assert np.abs(a[..., 2] - (g[2] - 1.7) * dx * dy).max() <= 1e-12 * dx * dy
    assert np.array_equal(a[..., :2], b[..., :2])


def pin_flux_forcing():
    """k-periodic uniform grid with a body (solid cells inside an IB layer),
    dpdz_set off: Calc_k_Flux is the fluid k flux (nvert(k-1) + nvert(k) <
    0.1) averaged over the planes and the mean area is dx dy times the fluid
    faces per plane (rhs2.c:1606-1627, bcs.c:262-278); the forcing adds
    (F - inlet_flux) dx dy / (dt A) to dP.z on every unmasked k face and
    nothing to the other components (momentum.c:394-397)."""
    Ni, Nj, Nk = 8, 9, 10
    dx, dy, dz = 0.3, 0.2, 0.25
    aff = Affine((Ni, Nj, Nk), np.diag([dx, dy, dz]), per=(1, 0, 1), immersed=1, dpdz_set=0)
    o = aff.o
    nv = np.zeros((aff.mz, aff.my, aff.mx))
    nv[3:8, 3:7, 2:6] = 1.0
    nv[4:7, 4:6, 3:5] = 3.0
    o.nvert[:] = nv.ravel()
    uc = np.random.default_rng(18).uniform(-1, 1, (aff.mz, aff.my, aff.mx, 3))
    F, A = o.calc_k_flux(np.ascontiguousarray(uc.ravel()))
    fluid = (np.roll(nv, 1, axis=0) + nv)[1:-1, 1:-1, 1:-1] < 0.1
    F_ref = uc[1:-1, 1:-1, 1:-1, 2][fluid].sum() / Nk
    A_ref = dx * dy * fluid.sum() / Nk
    assert abs(F - F_ref) <= 1e-13 * np.abs(uc[..., 2]).sum() and abs(A - A_ref) <= 1e-13 * A_ref
    p = np.zeros(o.P)
    dp = {}
    for fin in (F, F - 0.37):
        d = o.vec()
        o.pressure_gradient(p.copy(), d, kflux=(F, A), inlet_flux=fin)
        dp[fin] = d.reshape(aff.mz, aff.my, aff.mx, 3)
    assert not dp[F].any()   # at the target flux the forcing vanishes
    D = dp[F - 0.37]
    m = _unmasked(aff, nv, 2)
    ref = 0.37 * dx * dy / (aff.cfg["dt"] * A)
    assert m.sum() > 100 and np.abs(D[..., 2][m] - ref).max() <= 1e-13 * ref
    assert not D[..., 2][~m].any() and not D[..., :2].any()


def pin_ib_bc_face_rule():
    """IB_BC (momentum.c:2083-2124): every face of an IB cell (nvert == 1)
    gets the face-averaged velocity times its area vector, u(x_f).S for a
    linear u on the sheared grid; every other face keeps its flux."""
    rng = np.random.default_rng(12)
    G = rng.uniform(-1, 1, (3, 3))
    u0 = rng.uniform(-1, 1, 3)
    uf = lambda x: x @ G.T + u0  # noqa: E731
    big, small = Affine((16, 16, 16), immersed=1), Affine(immersed=1)
    for aff, nv in [(big, _sphere_mask(big))] + [(small, b) for b in small.bodies()[2:4]]:
        o = aff.o
        o.nvert[:] = nv.ravel()
        ucat = np.ascontiguousarray(aff.cells(uf).ravel())
        uc0 = rng.uniform(-1, 1, 3 * o.P)
        uc = uc0.copy()
        o.ib_bc(uc, ucat.copy())
        U, U0 = uc.reshape(aff.mz, aff.my, aff.mx, 3), uc0.reshape(aff.mz, aff.my, aff.mx, 3)
        inner = np.zeros(nv.shape, bool)
        inner[1:-1, 1:-1, 1:-1] = True
        ib = np.floor(nv) == 1
        for d in range(3):
            touched = inner & (ib | np.roll(ib, -1, axis=2 - d))
            ref = uf(aff.xf[d]) @ aff.S[d]
            assert touched.sum() > 10
            assert np.abs(U[..., d][touched] - ref[touched]).max() <= 1e-13 * np.abs(ref).max(), d
            assert np.array_equal(U[..., d][~touched], U0[..., d][~touched]), d


def pin_min_dt():
    """Calc_Minimum_dt (momentum.c:149-174) for a uniform stream on the
    sheared grid: the least time to cross a cell along xi, eta or zeta,
    min over m of V / |u.S_m|."""
    aff = Affine()
    o = aff.o
    for u in (np.array([0.9, -0.2, 0.35]), np.array([-0.1, 0.05, 1.3]), np.array([0.2, 1.1, -0.3])):
        uc = aff.face_flux(lambda x: np.broadcast_to(u, x.shape))
        ref = min(aff.vol / abs(u @ S) for S in aff.S)
        assert abs(o.min_dt(np.ascontiguousarray(uc.ravel())) - ref) <= 1e-13 * ref


def _ff2(o, ucont, ucat, nut=None, ren=None):
    if ren is not None:
        o.c.ren = ren
    rhs = o.vec()
    o.formfunction2(np.ascontiguousarray(ucont.ravel()).copy(), np.ascontiguousarray(ucat.ravel()),
                    o.sca() if nut is None else np.ascontiguousarray(nut.ravel()), rhs, 1.0)
    return rhs.reshape(o.mz, o.my, o.mx, 3)


def pin_ff2_affine_convective():
    """Linear velocity on the sheared walled grid, viscosity off (Re 1e300):
    rhs_f = -S_f.(G u + u tr G)(x_f) on every unmasked face, the wall rows
    included (their 2nd-order differences are exact for quadratic fluxes)."""
    rng = np.random.default_rng(13)
    aff = Affine(ren=1e300)
    G = rng.uniform(-1, 1, (3, 3))
    u0 = rng.uniform(-1, 1, 3)
    uf = lambda x: x @ G.T + u0  # noqa: E731
    R = _ff2(aff.o, aff.face_flux(uf), aff.cells(uf))
    nv = np.zeros((aff.mz, aff.my, aff.mx))
    for d in range(3):
        u = uf(aff.xf[d])
        ref = -((u @ G.T) + u * np.trace(G)) @ aff.S[d]
        m = _unmasked(aff, nv, d)
        assert np.abs(R[..., d][m] - ref[m]).max() <= 1e-12 * np.abs(ref).max(), d


def pin_ff2_affine_viscous():
    """Quadratic velocity on the sheared grid: the Re-dependent part of the
    right-hand side, (rhs(nu1) - rhs(nu2)) / (nu1 - nu2), equals
    S_f.(lap u + grad div u) on the faces away from the walls."""
    rng = np.random.default_rng(14)
    aff = Affine(N=(13, 12, 11))
    H = rng.uniform(-1, 1, (3, 3, 3))
    H = 0.5 * (H + H.transpose(0, 2, 1))
    G = rng.uniform(-1, 1, (3, 3))

    def uf(x):
        return 0.5 * np.einsum("abc,...b,...c->...a", H, x, x) + x @ G.T

    lap = np.einsum("abb->a", H)
    graddiv = np.einsum("bba->a", H)
    uc, u = aff.face_flux(uf), aff.cells(uf)
    R1 = _ff2(aff.o, uc, u, ren=1.0)
    R2 = _ff2(aff.o, uc, u, ren=2.0)
    D = (R1 - R2) / 0.5
    inner = (slice(3, -4), slice(3, -4), slice(3, -4))
    for d in range(3):
        ref = aff.S[d] @ (lap + graddiv)
        assert np.abs(D[inner][..., d] - ref).max() <= 1e-9 * abs(ref), d


def pin_ff2_skew_affine():
    """-skew with a linear velocity on the sheared walled grid, viscosity
    off: half the divergence form plus half the advective form,
    rhs_f = -S_f.(G u + u tr G / 2)(x_f), on every unmasked face, with the
    4th-order and the second-order scheme (the face differences and the
    interpolation to the cells are exact for linear fields, the wall rows'
    second-order ones included)."""
    rng = np.random.default_rng(15)
    G = rng.uniform(-1, 1, (3, 3))
    u0 = rng.uniform(-1, 1, 3)
    uf = lambda x: x @ G.T + u0  # noqa: E731
    for second in (0, 1):
        aff = Affine(ren=1e300)
        nv = np.zeros((aff.mz, aff.my, aff.mx))
        aff.o.c.skew = 1
        aff.o.c.second_order = second
        R = _ff2(aff.o, aff.face_flux(uf), aff.cells(uf))
        for d in range(3):
            u = uf(aff.xf[d])
            ref = -((u @ G.T) + 0.5 * u * np.trace(G)) @ aff.S[d]
            m = _unmasked(aff, nv, d)
            assert np.abs(R[..., d][m] - ref[m]).max() <= 1e-12 * np.abs(ref).max(), (second, d)


def _cell_lengths(M):
    """The cell length scales along x, y, z that the Clark term uses on an
    affine grid: |sum_m (V / |S_m|) e_m / |e_m||, the heights of the cell
    over its three face pairs along the unit lattice vectors."""
    e = [M[:, m] for m in range(3)]
    S = [np.cross(e[1], e[2]), np.cross(e[2], e[0]), np.cross(e[0], e[1])]
    V = float(np.linalg.det(M))
    return np.abs(sum(V / np.linalg.norm(S[m]) * e[m] / np.linalg.norm(e[m]) for m in range(3)))


def pin_ff2_clark_affine():
    """-clark with a quadratic velocity u = H x x / 2 + G x on the sheared
    grid: the face gradients are exact, so the model stress
    T_ab = sum_m du_a/dx_m du_b/dx_m dx_m^2 is quadratic along the lattice
    and the term's share of the right-hand side, rhs(clark) - rhs, equals
    -S_f.(div T)(x_f) / 12 on the faces away from the walls."""
    rng = np.random.default_rng(16)
    aff = Affine(N=(13, 12, 11))
    H = rng.uniform(-1, 1, (3, 3, 3))
    H = 0.5 * (H + H.transpose(0, 2, 1))
    G = rng.uniform(-1, 1, (3, 3))

    def uf(x):
        return 0.5 * np.einsum("abc,...b,...c->...a", H, x, x) + x @ G.T

    dl2 = _cell_lengths(aff.M) ** 2
    uc, u = aff.face_flux(uf), aff.cells(uf)
    R0 = _ff2(aff.o, uc, u)
    aff.o.c.clark = 1
    R1 = _ff2(aff.o, uc, u)
    D = R1 - R0
    inner = (slice(3, -4), slice(3, -4), slice(3, -4))
    for d in range(3):
        x = aff.xf[d][inner]
        J = np.einsum("abc,...c->...ab", H, x) + G                      # du_a / dx_b
        divT = np.einsum("amb,...bm,m->...a", H, J, dl2) + np.einsum("...am,bmb,m->...a", J, H, dl2)
        ref = -(divT @ aff.S[d]) / 12.0
        assert np.abs(D[inner][..., d] - ref).max() <= 1e-9 * np.abs(ref).max(), d


def _body_fields(aff, nv, uf, rng):
    """Cell velocity uf in fluid and IB cells, garbage in solid cells; face
    fluxes uf.S except on faces that touch a solid cell (garbage)."""
    u = aff.cells(uf).copy()
    solid = nv > 2.5
    u[solid] = rng.uniform(-1e3, 1e3, (solid.sum(), 3))
    uc = aff.face_flux(uf)
    for d in range(3):
        t = solid | np.roll(solid, -1, axis=2 - d)
        uc[..., d][t] = rng.uniform(-1e3, 1e3, t.sum())
    return uc, u


def pin_ff2_body_uniform_stream():
    """A uniform stream next to a planar body (layers of solid cells under
    an IB layer), garbage in the solid cells: rhs = 0 on every unmasked face
    (the QUICK branches of IB faces reproduce constants and never read a
    solid cell; the one-sided transverse derivatives of a constant vanish)."""
    rng = np.random.default_rng(15)
    aff = Affine(immersed=1)
    u0 = np.array([0.8, -0.45, 0.6])
    uf = lambda x: np.broadcast_to(u0, x.shape)  # noqa: E731
    for b, nv in enumerate(aff.bodies()):
        aff.o.nvert[:] = nv.ravel()
        uc, u = _body_fields(aff, nv, uf, rng)
        R = _ff2(aff.o, uc, u)
        for d in range(3):
            m = _unmasked(aff, nv, d)
            assert np.abs(R[..., d][m]).max() <= 1e-12 * np.abs(u0 @ aff.S[d]), (b, d)


def pin_ff2_body_viscous_linear():
    """A linear velocity next to a planar body on the sheared grid, garbage
    in the solid cells, and an eddy viscosity linear in the cell index t
    along the body's plane, garbage in IB, solid and ghost cells (each of the two directions of the
    plane in turn).  Every face flux a fluid cell sees
    is then nu_t(x_f) (G + G^T) S_f exactly: the one-sided transverse
    derivatives next to the body and the walls are exact for a linear
    field, and nu_t is taken from the fluid side at IB faces and from the
    interior cell at wall faces (momentum.c:885-897), which equals the face
    value because nu_t does not vary across those faces.  So the
    nu_t-dependent part of the right-hand side is S_f.(G + G^T) grad nu_t
    on every unmasked face, except next to the two walls normal to t."""
    rng = np.random.default_rng(16)
    aff = Affine(immersed=1, les=1)
    G = rng.uniform(-1, 1, (3, 3))
    T = G + G.T
    Minv = np.linalg.inv(aff.M)
    uf = lambda x: x @ G.T + 0.3  # noqa: E731
    for b, nv, t in [(b, nv, (b // 2 + s) % 3) for b, nv in enumerate(aff.bodies()) for s in (1, 2)]:
        aff.o.nvert[:] = nv.ravel()
        uc, u = _body_fields(aff, nv, uf, rng)
        it = aff.idx[..., t]
        nut = 0.7 + 0.3 * it
        bad = nv > 0.05
        bad[0], bad[-1], bad[:, 0], bad[:, -1], bad[:, :, 0], bad[:, :, -1] = (True,) * 6
        nut[bad] = rng.uniform(10, 1000, bad.sum())
        R1 = _ff2(aff.o, uc, u, nut=nut, ren=1e300)
        R0 = _ff2(aff.o, uc, u, nut=np.zeros(nv.shape), ren=1e300)
        grad = 0.3 * Minv[t]
        mt = [aff.mx, aff.my, aff.mz][t]
        edge = (it == 1) | (it == mt - 2)
        for d in range(3):
            m = _unmasked(aff, nv, d) & ~edge & ~np.roll(edge, -1, axis=2 - d)
            ref = aff.S[d] @ (T @ grad)
            assert m.sum() > 50
            assert np.abs((R1 - R0)[..., d][m] - ref).max() <= 1e-11 * abs(ref), (b, d)


def _rotation(a, b, c):
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def pin_rotation_invariance(name="C2"):
    """Rotating the curvilinear grid and the Cartesian velocity together
    leaves every contravariant quantity unchanged: Contra2Cart returns R u,
    and the eddy viscosity, the Formfunction_2 right-hand side, the pressure
    gradient, IB_BC's fluxes, the Poisson stencil, the divergence and the
    time step are the same (C2: stretched wavy grid, sphere, LES; C2w: the
    wall-function bottom as well).  A velocity component paired with the
    wrong metric component is not invariant."""
    R = _rotation(0.7, -0.4, 1.1)
    cfg, spec = W.make_config(name, (16, 12, 10))
    coor = W.grid(spec)
    o1 = Oracle(cfg, coor)
    if spec["immersed"]:
        o1.nvert_sphere(*W.SPHERE)
    o1.wall_mask()
    o2 = Oracle(cfg, np.ascontiguousarray((coor.reshape(-1, 3) @ R.T).ravel()))
    o2.nvert[:] = o1.nvert
    ucat, ucont = W.fields(cfg, spec, o1.g, o1.nvert)
    rot = lambda v: np.ascontiguousarray((v.reshape(-1, 3) @ R.T).ravel())  # noqa: E731
    u1, uc1 = ucat.copy(), ucont.copy()
    u2, uc2 = rot(ucat), ucont.copy()
    s1, s2 = o1.sca(), o2.sca()
    o1.contra2cart_ustar(uc1, u1, s1)
    o2.contra2cart_ustar(uc2, u2, s2)
    assert rel_err(u2, rot(u1)) < 1e-12 and rel_err(s2, s1) < 1e-12
    n1, n2 = o1.sca(), o2.sca()
    o1.eddy_viscosity(u1, W.CS_LES, n1)
    o2.eddy_viscosity(u2, W.CS_LES, n2)
    assert rel_err(n2, n1) < 1e-11
    r1, r2 = o1.vec(), o2.vec()
    o1.formfunction2(uc1.copy(), u1, n1, r1, 1.0)
    o2.formfunction2(uc2.copy(), u2, n2, r2, 1.0)
    assert rel_err(r2, r1) < 1e-11
    p = W.random_vector(cfg, 1.0, 17)[: o1.P].copy()
    d1, d2 = o1.vec(), o2.vec()
    o1.pressure_gradient(p.copy(), d1)
    o2.pressure_gradient(p.copy(), d2)
    assert rel_err(d2, d1) < 1e-11
    a1, a2 = uc1.copy(), uc2.copy()
    b1, b2 = u1.copy(), u2.copy()
    o1.ib_bc(a1, b1)
    o2.ib_bc(a2, b2)
    assert rel_err(a2, a1) < 1e-12 and rel_err(b2, rot(b1)) < 1e-12
    assert rel_err(o2.poisson_stencil(), o1.poisson_stencil()) < 1e-12
    m1, _ = o1.divergence(uc1)
    m2, _ = o2.divergence(uc2)
    assert abs(m2 - m1) <= 1e-12 * m1
    assert abs(o2.min_dt(uc2) - o1.min_dt(uc1)) <= 1e-12 * o1.min_dt(uc1)


# ---------------------------------------------------------------------------
# the 2-D Taylor-Green vortex (the reference's known answer)
# ---------------------------------------------------------------------------
TG_A = 2.0 * math.pi


def tg_grid(N, eps):
    """Unit-periodic x-y grid x = xi + eps sin(2 pi eta), y = eta + eps
    sin(2 pi xi), z uniform over 5 cells; periodic in all directions."""
    Nk = 5
    mx, my, mz = N + 2, N + 2, Nk + 2
    kk, jj, ii = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    xi, et = ii / N, jj / N
    coor = np.stack([xi + eps * np.sin(2 * np.pi * et), et + eps * np.sin(2 * np.pi * xi), kk * 0.05], -1)
    cfg = dict(mx=mx, my=my, mz=mz, i_periodic=1, j_periodic=1, k_periodic=1, bctype=[7] * 6, immersed=0, les=0,
               second_order=0, dpdz_set=1, weno=0, inviscid=0, ren=10.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    return cfg, np.ascontiguousarray(coor.ravel())


def tg_u(x):
    u = -np.cos(TG_A * x[..., 0]) * np.sin(TG_A * x[..., 1])
    v = np.sin(TG_A * x[..., 0]) * np.cos(TG_A * x[..., 1])
    return np.stack([u, v, np.zeros_like(u)], -1)


def tg_p(x):
    return -0.25 * (np.cos(2 * TG_A * x[..., 0]) + np.cos(2 * TG_A * x[..., 1]))


def _periodic_fill(a):
    a[:, :, 0], a[:, :, -1] = a[:, :, -2], a[:, :, 1]
    a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
    a[0], a[-1] = a[-2], a[1]
    return a


def tg_case(N, eps):
    cfg, coor = tg_grid(N, eps)
    o = Oracle(cfg, coor)
    X = coor.reshape(cfg["mz"], cfg["my"], cfg["mx"], 3)
    cent = o.g["cent"].reshape(cfg["mz"], cfg["my"], cfg["mx"], 3)
    u = _periodic_fill(tg_u(cent).copy())
    p = _periodic_fill(tg_p(cent).copy())
    # face centres (mean of the four face nodes, bcs.c:3240-3270) and fluxes
    fc = np.zeros((3,) + X.shape)
    fc[0, 1:, 1:, :] = 0.25 * (X[1:, 1:, :] + X[:-1, 1:, :] + X[1:, :-1, :] + X[:-1, :-1, :])
    fc[1, 1:, :, 1:] = 0.25 * (X[1:, :, 1:] + X[:-1, :, 1:] + X[1:, :, :-1] + X[:-1, :, :-1])
    fc[2, :, 1:, 1:] = 0.25 * (X[:, 1:, 1:] + X[:, :-1, 1:] + X[:, 1:, :-1] + X[:, :-1, :-1])
    S = [o.g[n].reshape(X.shape) for n in ("icsi", "jeta", "kzet")]
    uf = [tg_u(fc[d]) for d in range(3)]
    uc = np.stack([np.einsum("...c,...c->...", uf[d], S[d]) for d in range(3)], -1)
    return cfg, o, u, p, uc, uf, S


def tg_rhs_error(N, eps):
    """max over interior faces of |Formfunction_2 - Pressure_Gradient -
    S.du/dt| relative to max |S.du/dt|: at t = 0 the exact solution has
    du/dt = -2 nu a^2 u (the decay rate in rhs2.c:1386-1388)."""
    cfg, o, u, p, uc, uf, S = tg_case(N, eps)
    rhs = o.vec()
    o.formfunction2(np.ascontiguousarray(uc.ravel()), np.ascontiguousarray(u.ravel()), o.sca(), rhs, 1.0)
    dp = o.vec()
    o.pressure_gradient(np.ascontiguousarray(p.ravel()), dp)
    got = (rhs - dp).reshape(uc.shape)[1:-1, 1:-1, 1:-1]
    nu = 1.0 / cfg["ren"]
    ref = np.stack([np.einsum("...c,...c->...", -2 * nu * TG_A ** 2 * uf[d], S[d]) for d in range(3)],
                   -1)[1:-1, 1:-1, 1:-1]
    return np.abs(got - ref).max() / np.abs(ref).max()


def tg_projection_error(N, eps):
    """Project the exact face fluxes (whose discrete divergence is the
    midpoint rule's error) and measure how far they move."""
    cfg, o, u, p, uc, uf, S = tg_case(N, eps)
    U = np.ascontiguousarray(uc.ravel())
    b = o.sca()
    o.poisson_rhs(U, b)
    phi = o.sca()
    it, rel = o.poisson_solve(b, phi, rtol=1e-12)
    assert rel <= 1e-12
    U1 = U.copy()
    o.projection(phi, U1)
    d1, _ = o.divergence(U1)
    assert d1 <= 1e-13
    inner = lambda a: a.reshape(uc.shape)[1:-1, 1:-1, 1:-1, :2]  # noqa: E731
    return np.abs(inner(U1) - inner(U)).max() / np.abs(inner(U)).max()


def pin_taylor_green():
    """Second-order convergence of the momentum right-hand side minus the
    pressure gradient, and of the projection, to the Taylor-Green answer on
    a uniform and on a curvilinear periodic grid."""
    for eps in (0.0, 0.03):
        e = [tg_rhs_error(N, eps) for N in (16, 32, 64)]   # measured 3.1e-2 7.5e-3 1.9e-3 / 4.7e-2 1.3e-2 3.3e-3
        assert e[2] < 0.005, (eps, e)
        assert e[0] / e[1] > 3.4 and e[1] / e[2] > 3.6, (eps, e)
    # on the uniform grid the midpoint fluxes of the vortex are discretely
    # solenoidal (the sine differences cancel), so the projection moves
    # nothing; on the curved grid it moves them by O(h^2)
    assert tg_projection_error(16, 0.0) < 1e-14
    e = [tg_projection_error(N, 0.03) for N in (16, 32)]        # measured 1.7e-3, 4.1e-4
    assert e[0] / e[1] > 3.6 and e[1] < 1e-3, e


PINS = {f.__name__: f for f in (pin_pressure_gradient_affine_with_body, pin_mean_pressure_gradient, pin_flux_forcing,
                                pin_ib_bc_face_rule,
                                pin_min_dt, pin_ff2_affine_convective, pin_ff2_affine_viscous,
                                pin_ff2_skew_affine, pin_ff2_clark_affine,
                                pin_ff2_body_uniform_stream, pin_ff2_body_viscous_linear, pin_rotation_invariance,
                                pin_taylor_green)}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


def test_rotation_invariance_wall_function():
    pin_rotation_invariance("C2w")


# ---------------------------------------------------------------------------
# planted slips: each mutant of the oracle must fail at least one pin
# ---------------------------------------------------------------------------
# (name, [(old, new, count)], pins expected to catch it)
MUTANTS = [
    ("viscous transpose term transposed",
     [("r21 * nrm1 + r31 * nrm2) * ajc * (nu)", "r12 * nrm1 + r13 * nrm2) * ajc * (nu)", 1)],
     ["pin_ff2_affine_viscous", "pin_rotation_invariance", "pin_taylor_green"]),
    ("viscous cross-metric g31 term dropped",
     [("g31 * dz[0] + r11 * nrm0 + r21 * nrm1 + r31 * nrm2) * ajc * (nu)", "r11 * nrm0 + r21 * nrm1 + r31 * nrm2) * ajc * (nu)", 1)],
     ["pin_ff2_affine_viscous", "pin_taylor_green"]),
    ("one-sided transverse derivative (body above) at quarter weight",
     [("de[q] = (U3(k, j, i + 1, q) + U3(k, j, i, q) - U3(k, j - 1, i + 1, q) - U3(k, j - 1, i, q)) * 0.5;",
       "de[q] = (U3(k, j, i + 1, q) + U3(k, j, i, q) - U3(k, j - 1, i + 1, q) - U3(k, j - 1, i, q)) * 0.25;", 1)],
     ["pin_ff2_body_viscous_linear"]),
    ("one-sided transverse derivative (body below) at quarter weight",
     [("de[q] = (U3(k, j + 1, i + 1, q) + U3(k, j + 1, i, q) - U3(k, j, i + 1, q) - U3(k, j, i, q)) * 0.5;",
       "de[q] = (U3(k, j + 1, i + 1, q) + U3(k, j + 1, i, q) - U3(k, j, i + 1, q) - U3(k, j, i, q)) * 0.25;", 1)],
     ["pin_ff2_body_viscous_linear"]),
    ("QUICK at an IB face reads the cell behind the body",
     [("um * (0.125 * (-UB(n + 2, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q))",
       "um * (0.125 * (-UB(n - 1, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q))", 1)],
     ["pin_ff2_body_uniform_stream"]),
    ("QUICK weight",
     [("up * (0.125 * (-UB(n - 1, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q))",
       "up * (0.125 * (-UB(n - 1, q) - 2. * UB(n, q) + 2. * UB(n + 1, q)) + UB(n, q))", 1)],
     ["pin_ff2_body_uniform_stream"]),
    ("4th-order interpolation weight",
     [("0.0625 * (-UB(L, q) + 9. * UB(n, q) + 9. * UB(n + 1, q) - UB(R, q))",
       "0.0625 * (-UB(L, q) + 9. * UB(n, q) + 9. * UB(n + 1, q) - UB(n, q))", 1)],
     ["pin_ff2_affine_convective", "pin_taylor_green"]),
    ("4th-order divergence weight",
     [("f[q] = 9. / 8. * dv[q] + -1. / 8. * d4[q];", "f[q] = 9. / 8. * dv[q] + -1. / 6. * d4[q];", 1)],
     ["pin_ff2_affine_convective", "pin_taylor_green"]),
    ("eddy viscosity of an IB face from the body side",
     [("if ((n == 0 && !per) || NB(n) > 0.1) nut = dir == 0 ? nu_t[IX(k, j, i + 1)]",
       "if ((n == 0 && !per) || NB(n) > 0.1) nut = dir == 0 ? nu_t[id]", 1)],
     ["pin_ff2_body_viscous_linear"]),
    ("pressure gradient: one-sided difference at quarter weight",
     [("dpde = (PP(k, j, i) - PP(k, j - 1, i) + PP(k, j, i + 1) - PP(k, j - 1, i + 1)) * 0.5;",
       "dpde = (PP(k, j, i) - PP(k, j - 1, i) + PP(k, j, i + 1) - PP(k, j - 1, i + 1)) * 0.25;", 1)],
     ["pin_pressure_gradient_affine_with_body"]),
    ("pressure gradient: cross metric g12 from zeta",
     [("double g12_i = (ie[0] * ic[0]", "double g12_i = (iz[0] * ic[0]", 1)],
     ["pin_pressure_gradient_affine_with_body", "pin_taylor_green"]),
    ("pressure gradient: one-sided branch reads the IB cell",
     [("dpdz = (PP(k + 1, j, i) - PP(k, j, i) + PP(k + 1, j + 1, i) - PP(k, j + 1, i)) * 0.5;",
       "dpdz = (PP(k + 1, j, i) - PP(k - 1, j, i) + PP(k + 1, j + 1, i) - PP(k - 1, j + 1, i)) * 0.25;", 1)],
     ["pin_pressure_gradient_affine_with_body"]),
    ("mean pressure gradient with the wrong sign",
     [("dpdz += c->mean_pressure_gradient * dzz;", "dpdz -= c->mean_pressure_gradient * dzz;", 1)],
     ["pin_mean_pressure_gradient"]),
    ("flux forcing with the wrong sign",
     [("dpdz += dzz * (mean_k_flux - inlet_flux) / c->dt / mean_k_area;",
       "dpdz -= dzz * (mean_k_flux - inlet_flux) / c->dt / mean_k_area;", 1)],
     ["pin_flux_forcing"]),
    ("Calc_k_Flux not averaged over the planes",
     [("*mean_k_flux = f / (mz - 2);", "*mean_k_flux = f;", 1)],
     ["pin_flux_forcing"]),
    ("Calc_k_Flux fluid test on one cell",
     [("if (nvert[IX(k - 1, j, i)] + nvert[IX(k, j, i)] < 0.1) {\n                    f +=",
       "if (nvert[IX(k, j, i)] < 0.1) {\n                    f +=", 1)],
     ["pin_flux_forcing"]),
    ("IB_BC j face from the cell below",
     [("if ((int)(NV(k, j + 1, i)) == 1) {\n                        ucx = (U3(k, j + 1, i, 0) + U3(k, j, i, 0)) * 0.5;",
       "if ((int)(NV(k, j + 1, i)) == 1) {\n                        ucx = (U3(k, j - 1, i, 0) + U3(k, j, i, 0)) * 0.5;", 1)],
     ["pin_ib_bc_face_rule"]),
    ("IB_BC i face with the eta metric",
     [("VX(ucont, id) = (ucx * ic[0] + ucy * ic[1] + ucz * ic[2]) * f;\n                    }\n                    if ((int)(NV(k, j + 1, i)) == 1)",
       "VX(ucont, id) = (ucx * je[0] + ucy * je[1] + ucz * je[2]) * f;\n                    }\n                    if ((int)(NV(k, j + 1, i)) == 1)", 1)],
     ["pin_ib_bc_face_rule", "pin_rotation_invariance"]),
    ("Calc_Minimum_dt pairs the eta flux with the wrong Jacobian index",
     [("a = fabs(1. / VY(ucont, id) / g->aj[id]);", "a = fabs(1. / VX(ucont, id) / g->aj[id]);", 1)],
     ["pin_min_dt"]),
    ("skew: divergence form not halved",
     [("d3v[q] *= 0.5;", "d3v[q] *= 1.0;", 1)],
     ["pin_ff2_skew_affine"]),
    ("skew: 4th-order velocity difference reversed",
     [("const double d4 = (UB(R, q) - UB(L, q)) * (1. / denom);",
       "const double d4 = (UB(L, q) - UB(R, q)) * (1. / denom);", 1)],
     ["pin_ff2_skew_affine"]),
    ("skew: wall faces keep the divisor 3",
     [("else L = n, R = n + 1, denom = 1.;", "else L = n, R = n + 1;", 1)],
     ["pin_ff2_skew_affine"]),
    ("skew: interpolation to the cell from the wrong side",
     [("f[q] += 9. / 8. * 0.5 * (adv1[3 * id + q] + adv1[3 * im + q])",
       "f[q] += 9. / 8. * 0.5 * (adv1[3 * id + q] + adv1[3 * a1 + q])", 1)],
     ["pin_ff2_skew_affine"]),
    ("clark: the y length scale taken from x",
     [("const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dzl * dzl;",
       "const double dx2 = dx * dx, dy2 = dx * dx, dz2 = dzl * dzl;", 1)],
     ["pin_ff2_clark_affine"]),
    ("clark: t23 from the u gradient",
     [("const double t23 = (dv_dx * dw_dx * dx2", "const double t23 = (du_dx * dw_dx * dx2", 1)],
     ["pin_ff2_clark_affine"]),
    ("clark: the stress applied with the wrong sign",
     [("v0 -= (t11 * nrm0 + t12 * nrm1 + t13 * nrm2) / 12.;", "v0 += (t11 * nrm0 + t12 * nrm1 + t13 * nrm2) / 12.;", 1)],
     ["pin_ff2_clark_affine"]),
]


def _build_mutant(d, n, edits):
    src = open(os.path.join(OP.ORACLE_DIR, "vfs_oracle.c")).read()
    for old, new, count in edits:
        assert src.count(old) == count, (n, old)
        src = src.replace(old, new)
    sub = os.path.join(d, f"m{n}")
    os.makedirs(sub)
    shutil.copy(os.path.join(OP.ORACLE_DIR, "vfs_oracle.h"), sub)
    with open(os.path.join(sub, "vfs_oracle.c"), "w") as f:
        f.write(src)
    so = os.path.join(sub, "liboracle.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared", "-o", so,
                    os.path.join(sub, "vfs_oracle.c"), "-lm"], check=True, capture_output=True)
    return so


def test_planted_slips_fail_a_pin():
    from concurrent.futures import ThreadPoolExecutor
    d = tempfile.mkdtemp(prefix="vfs_mut_")
    try:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
            libs = list(ex.map(lambda a: _build_mutant(d, a[0], a[1][1]), enumerate(MUTANTS)))
        missed = []
        for (name, edits, pins), so in zip(MUTANTS, libs):
            caught = []
            with OP.using(so):
                for p in pins:
                    try:
                        PINS[p]()
                    except AssertionError:
                        caught.append(p)
            if not caught:
                missed.append(name)
        assert not missed, f"slips no pin caught: {missed}"
    finally:
        shutil.rmtree(d, ignore_errors=True)
