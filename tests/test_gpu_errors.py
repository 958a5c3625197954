"""Failure modes of the drop-in path reported as return codes, where the
reference prints and carries on or exits (momentum.c:157 "Negative
jacobian"; poisson_hypre.c:581-597, whose HYPRE solve reports nothing).

- vfs_build_metrics: -7 for a folded grid (non-positive Jacobian);
- vfs_poisson_setup: -8 when no cell is fluid;
- vfs_poisson_solve: 1 when maxit is reached above rtol (the last iterate is
  returned), -6 for a non-finite right-hand side, also one that a NaN
  velocity field produces through VolumeFlux and PoissonRHS.
"""
import sys

import numpy as np
import pytest

from cases import gpu_case, oracle_case
from vfsgpu import workloads as W
import vfsgpu as V

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _small(name="C2", N=(24, 16, 12)):
    cfg, spec = W.make_config(name, N)
    coor = W.grid(spec)
    return cfg, spec, coor


def test_folded_grid_is_rejected():
    say("[errors] folded grid, C1 16x12x14")
    cfg, spec, coor = _small("C1", (16, 12, 14))
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))   # the plain grid is accepted
    c = coor.reshape(-1, 3).copy()
    c[:, 0] = -c[:, 0]                                  # mirrored in x: every Jacobian < 0
    with pytest.raises(V.VfsError) as e:
        ctx.build_metrics(V.DeviceArray.from_host(c.ravel()))
    assert e.value.rc == -7
    assert "Jacobian" in str(e.value)
    # one folded cell: node plane i = 5 moved beyond plane i = 6
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    c = coor.reshape(mz, my, mx, 3).copy()
    c[:, :, 5, 0] = c[:, :, 6, 0] + 0.5 * (c[:, :, 6, 0] - c[:, :, 4, 0])
    with pytest.raises(V.VfsError) as e:
        ctx.build_metrics(V.DeviceArray.from_host(c.ravel()))
    assert e.value.rc == -7


def test_empty_fluid_set_is_rejected():
    say("[errors] all-solid mask, C1 16x12x14")
    cfg, spec, coor = _small("C1", (16, 12, 14))
    cfg["immersed"] = 1
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(np.full(ctx.P, 3.0)))
    with pytest.raises(V.VfsError) as e:
        ctx.poisson_setup()
    assert e.value.rc == -8


@pytest.fixture(scope="module")
def c2():
    say("[errors] C2 24x16x12")
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (24, 16, 12))
    ctx = gpu_case(cfg, spec, coor)
    ctx.poisson_setup()
    b = ctx.sca()
    duc = V.DeviceArray.from_host(ucont)
    ctx.volume_flux(duc)
    ctx.poisson_rhs(duc, b)
    return dict(ctx=ctx, b=b, ucont=ucont)


@pytest.mark.parametrize("precision", [0, 1])
def test_not_converged_returns_one(c2, precision):
    ctx = c2["ctx"]
    phi = ctx.sca()
    with pytest.raises(V.VfsNotConverged) as e:
        ctx.poisson_solve(c2["b"], phi, rtol=1e-12, precision=precision, maxit=1)
    assert e.value.rc == 1
    assert e.value.iters >= 1
    assert 1e-12 < e.value.relres < 1.0
    # the last iterate came back, and relres is its FP64 residual
    y = ctx.sca()
    ctx.poisson_apply(phi, y)
    bh, yh = c2["b"].host(), y.host()
    gid = ctx.poisson_gid().host()
    r = np.linalg.norm((bh - yh)[gid >= 0]) / np.linalg.norm(bh[gid >= 0])
    assert abs(r - e.value.relres) <= 1e-6 * e.value.relres
    it, rel = ctx.poisson_solve(c2["b"], phi, rtol=1e-8, precision=precision)
    assert rel <= 1e-8


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_rhs_returns_minus_six(c2, bad):
    ctx = c2["ctx"]
    bh = c2["b"].host()
    gid = ctx.poisson_gid().host().reshape(ctx.mz, ctx.my, ctx.mx)
    inner = np.zeros(gid.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True   # gid also numbers the periodic ghost copies
    cells = np.flatnonzero((gid >= 0) & inner)
    bh[int(cells[len(cells) // 2])] = bad
    with pytest.raises(V.VfsError) as e:
        ctx.poisson_solve(V.DeviceArray.from_host(bh), ctx.sca(), rtol=1e-8)
    assert e.value.rc == -6


def test_nan_velocity_reaches_the_solve_as_minus_six(c2):
    ctx = c2["ctx"]
    uc = c2["ucont"].copy()
    mx, my = ctx.mx, ctx.my
    uc[3 * ((5 * my + 6) * mx + 7)] = np.nan
    duc = V.DeviceArray.from_host(uc)
    b = ctx.sca()
    ctx.volume_flux(duc)
    ctx.poisson_rhs(duc, b)
    with pytest.raises(V.VfsError) as e:
        ctx.poisson_solve(b, ctx.sca(), rtol=1e-8, precision=1)
    assert e.value.rc == -6
