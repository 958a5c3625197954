"""Derivatives of the implicit momentum residual F(U) (FormFunction_SNES,
momentum.c:2197-2301) on the GPU, and the Newton-Krylov solve built on them
(Implicit_MatrixFree, implicitsolver.c:4196-4441).

What pins them:
- JVP: with nu_t and the IB-node ucat frozen, F is quadratic in U (the
  central convective flux is a product of two linear forms of U, the
  viscous flux is linear) and piecewise quadratic where the IB faces upwind
  with |U| (momentum.c:731-732).  A central difference of the oracle's F is
  therefore exact up to rounding, so J w must match it to ~1e-9.
- VJP: the adjoint identity <v, J w> = <J^T v, w> to 1e-12 of |v| |J w|
  for random v, w, single entries against the JVP's, and the same bits on
  repeated calls.
- Newton-Krylov: the oracle's F at the returned U is below the requested
  relative tolerance (the reference's PETSc iterates are not reproducible
  without PETSc 3.1: the converged state is the check).
- torch.autograd.gradcheck through the C-ABI (own process).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from cases import gpu_case, oracle_case
from conftest import rel_err
from vfsgpu import workloads as W
import vfsgpu as V

pytestmark = pytest.mark.gpu
CASES = ["C1", "C2", "C2-weno"]
DT = 0.05     # keeps the FF2 part of J comparable to the -1/dt identity part


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", params=CASES)
def dcase(request):
    name = request.param
    say(f"[derivatives] {name}")
    weno = name.endswith("-weno")
    name = name.split("-")[0]
    cfg, spec, coor, o, ucat, ucont = oracle_case(name, dt=DT)
    if weno:
        cfg = dict(cfg, weno=1)
        o = __import__("oracle_py").Oracle(cfg, coor)
        if spec["immersed"]:
            o.nvert_sphere(*W.SPHERE)
    ctx = gpu_case(cfg, spec, coor)
    uc, u = ucont.copy(), ucat.copy()
    o.contra2cart(uc, u)
    nut = o.sca()
    if cfg["les"]:
        o.eddy_viscosity(u, W.CS_LES, nut)
    U = uc + 0.01 * np.abs(uc).max() * W.random_vector(cfg, 1.0, 61)
    Uo = uc.copy()
    rhs_o = 0.1 * np.abs(uc).max() * W.random_vector(cfg, 1.0, 62)
    dP = 0.1 * np.abs(uc).max() * W.random_vector(cfg, 1.0, 63)
    return dict(name=name, weno=weno, cfg=cfg, o=o, ctx=ctx, U=U, Uo=Uo, rhs_o=rhs_o, dP=dP, ucat=u, nut=nut)


def oracle_F(c, U):
    o = c["o"]
    lu, uu, r = o.vec(), c["ucat"].copy(), o.vec()
    o.formfunction_snes(U, c["Uo"], c["rhs_o"], c["dP"], c["nut"], lu, uu, r)
    return r


def dev(c):
    return [V.DeviceArray.from_host(x) for x in (c["U"], c["ucat"], c["nut"])]


@pytest.mark.parametrize("seed", [14001, 14002])
def test_jvp_matches_central_difference(dcase, seed):
    c = dcase
    w = W.unit_vector(c["cfg"], seed) * np.sqrt(c["U"].size) * np.abs(c["U"]).max()
    h = 1e-7 if c["weno"] else 1e-6
    fd = (oracle_F(c, c["U"] + h * w) - oracle_F(c, c["U"] - h * w)) / (2 * h)
    dU, du, dn = dev(c)
    jw = c["ctx"].vec()
    c["ctx"].rhs_jvp(dU, du, dn, V.DeviceArray.from_host(w), jw)
    err = rel_err(jw.host(), fd)
    say(f"  jvp vs central difference: {err:.2e}")
    # quadratic F: exact up to rounding; with WENO weights F is smooth but not
    # quadratic, so the central difference carries an O(h^2) error
    assert err < (1e-6 if c["weno"] else 1e-8)


def test_jvp_is_linear(dcase):
    c = dcase
    cfg = c["cfg"]
    w1, w2 = W.random_vector(cfg, 1.0, 71), W.random_vector(cfg, 1.0, 72)
    dU, du, dn = dev(c)
    out = []
    for w in (w1, w2, 2.0 * w1 - 3.0 * w2):
        jw = c["ctx"].vec()
        c["ctx"].rhs_jvp(dU, du, dn, V.DeviceArray.from_host(w), jw)
        out.append(jw.host())
    assert rel_err(out[2], 2.0 * out[0] - 3.0 * out[1]) < 1e-13


@pytest.mark.parametrize("seeds", [(14001, 24001), (14003, 24003), (14005, 24005)])
def test_adjoint_identity(dcase, seeds):
    c = dcase
    cfg = c["cfg"]
    w = W.unit_vector(cfg, seeds[0])
    v = W.unit_vector(cfg, seeds[1])
    dU, du, dn = dev(c)
    jw, jtv = c["ctx"].vec(), c["ctx"].vec()
    c["ctx"].rhs_jvp(dU, du, dn, V.DeviceArray.from_host(w), jw)
    c["ctx"].rhs_vjp(dU, du, dn, V.DeviceArray.from_host(v), jtv)
    Jw, JTv = jw.host(), jtv.host()
    lhs, rhs = float(v @ Jw), float(JTv @ w)
    scale = np.linalg.norm(v) * np.linalg.norm(Jw)
    say(f"  <v,Jw> = {lhs:.15e}  <J^T v,w> = {rhs:.15e}  rel {abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= 1e-12 * scale


def test_vjp_columns_against_jvp(dcase):
    """(J^T e_r)_c = (J e_c)_r on single entries near the walls and the body."""
    c = dcase
    cfg = c["cfg"]
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    dU, du, dn = dev(c)
    picks = [(1, 1, 1, 0), (mx // 2, my // 2, mz // 2, 1), (mx - 2, my - 2, mz - 2, 2), (2, my // 2 + 1, 3, 1)]
    n = c["U"].size
    for (i, j, k, q) in picks:
        r = 3 * ((k * my + j) * mx + i) + q
        e = np.zeros(n)
        e[r] = 1.0
        jtv = c["ctx"].vec()
        c["ctx"].rhs_vjp(dU, du, dn, V.DeviceArray.from_host(e), jtv)
        row = jtv.host()
        scale = np.abs(row).max()
        cols = np.argsort(-np.abs(row))[:12]
        for col in cols:
            ec = np.zeros(n)
            ec[col] = 1.0
            jw = c["ctx"].vec()
            c["ctx"].rhs_jvp(dU, du, dn, V.DeviceArray.from_host(ec), jw)
            a, b = jw.host()[r], row[col]
            assert abs(a - b) <= 1e-12 * scale, (r, col, a, b)


def test_vjp_is_deterministic(dcase):
    """J^T v gathers every adjoint element in a fixed order (no atomics): calls
    on the same inputs give the same bits."""
    c = dcase
    v = V.DeviceArray.from_host(W.unit_vector(c["cfg"], 24007))
    dU, du, dn = dev(c)
    out = []
    for _ in range(3):
        jtv = c["ctx"].vec()
        c["ctx"].rhs_vjp(dU, du, dn, v, jtv)
        out.append(jtv.host())
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_newton_krylov_converges(dcase):
    """Implicit step at CFL 2 (dt = 2 x Calc_Minimum_dt, momentum.c:153-174)."""
    c = dcase
    o = c["o"]
    Uo = c["Uo"]
    dt = 2.0 * o.min_dt(Uo)
    cfg2, spec, coor, o2, _, _ = oracle_case(c["name"], dt=dt)
    if c["weno"]:
        cfg2 = dict(cfg2, weno=1)
        o2 = __import__("oracle_py").Oracle(cfg2, coor)
        if spec["immersed"]:
            o2.nvert_sphere(*W.SPHERE)
    ctx = gpu_case(cfg2, spec, coor)
    # rhs_o = FF2(U_o) as left by the previous step, dP from a smooth pressure
    uc, u = Uo.copy(), c["ucat"].copy()
    rhs_o = o2.vec()
    o2.formfunction2(uc, u, c["nut"], rhs_o, 1.0)
    p = o2.sca()
    cent = o2.g["cent"].reshape(-1, 3)
    p[:] = 0.01 * np.sin(cent[:, 0]) * np.cos(np.pi * cent[:, 1] / 2)
    dP = o2.vec()
    o2.pressure_gradient(p.copy(), dP)
    U = Uo.copy()

    def F(Ux):
        lu, uu, r = o2.vec(), c["ucat"].copy(), o2.vec()
        o2.formfunction_snes(Ux, Uo, rhs_o, dP, c["nut"], lu, uu, r)
        return r

    f0 = np.linalg.norm(F(U))
    dU = V.DeviceArray.from_host(U)
    du = V.DeviceArray.from_host(c["ucat"])
    args = [V.DeviceArray.from_host(x) for x in (Uo, rhs_o, dP, c["nut"])]
    for rtol in (1e-4, 1e-8):
        ni, ki, fn = ctx.implicit_matrixfree(dU, *args, du, rtol=rtol, max_newton=20, restart=30, max_krylov=1000)
        fx = np.linalg.norm(F(dU.host()))
        say(f"  CFL 2 dt={dt:.3e} rtol={rtol:g}: newton {ni} krylov {ki} |F|/|F0| gpu {fn / f0:.2e} oracle {fx / f0:.2e}")
        assert fx <= rtol * f0 * 1.01
        assert 0 < ni <= (15 if c["weno"] else 10)
    ctx.close()


def test_torch_gradcheck():
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "gradcheck_snes.py")], capture_output=True, text=True,
                       timeout=110)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "gradcheck ok" in r.stdout
