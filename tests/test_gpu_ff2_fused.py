"""The LDS-staged Formfunction_2 path against the reference-structured
passes (face arrays, cell divergence, projection; momentum.c:672-1900) on the
GPU: bit-identical right-hand sides.

The sizes put tile edges (64 cells in i, 8 in j), k-chunk edges (32 planes)
and periodic seams inside the domain, with and without the immersed sphere,
wall-function walls and the k inlet / outlet; the scheme variants (WENO3, the
inviscid path, second order) take the same fields.  A mask crossing the tile
edges makes both the mask-free and the general face paths run in one launch.
"""
import sys

import numpy as np
import pytest

from cases import gpu_case
from vfsgpu import workloads as W
import vfsgpu as V

pytestmark = pytest.mark.gpu

SIZES = {
    "C1": (130, 18, 70),    # periodic i and k, two full i tiles and a ragged third
    "C2": (70, 26, 66),     # the sphere across tile and chunk edges
    "C2w": (66, 17, 33),    # wall-function bottom
    "C1o": (65, 9, 35),     # k inlet / outlet
    "C1wo": (5, 5, 5),      # fewer cells than a tile in every direction
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(ctx, cfg, uc, u, nut, rhs0, monkeypatch, path):
    if path:
        monkeypatch.setenv("VFS_FF2_PATH", path)
    else:
        monkeypatch.delenv("VFS_FF2_PATH", raising=False)
    duc, du, dn, dr = (V.DeviceArray.from_host(x) for x in (uc, u, nut, rhs0))
    ctx.formfunction2(duc, du, dn, dr, 0.5)
    return dr.host(), duc.host()


@pytest.mark.parametrize("scheme", [dict(), dict(weno=1), dict(inviscid=1), dict(second_order=1)],
                         ids=["4th", "weno", "inviscid", "2nd"])
@pytest.mark.parametrize("name", list(SIZES))
def test_fused_equals_passes(name, scheme, monkeypatch):
    N = SIZES[name]
    say(f"[ff2-fused] {name} {N} {scheme}")
    cfg, spec = W.make_config(name, N)
    cfg.update(scheme)
    coor = W.grid(spec)
    ctx = gpu_case(cfg, spec, coor)
    nvert = ctx.nvert().host()
    metrics = {m: ctx.metric(m).host() for m in ("cent", "icsi", "jeta", "kzet")}
    ucat, ucont = W.fields(cfg, spec, metrics, nvert)
    # ucat from the contravariant fluxes, nu_t from the Smagorinsky model
    duc, du = V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat)
    ctx.contra2cart(duc, du)
    u = du.host()
    nut = np.zeros_like(nvert)
    if cfg["les"]:
        dn = ctx.sca()
        ctx.eddy_viscosity(du, W.CS_LES, dn)
        nut = dn.host()
    rhs0 = W.random_vector(cfg, 0.3, 78)
    r_pass, uc_pass = _run(ctx, cfg, ucont, u, nut, rhs0, monkeypatch, "passes")
    r_fused, uc_fused = _run(ctx, cfg, ucont, u, nut, rhs0, monkeypatch, None)
    assert np.isfinite(r_fused).all()
    diff = np.abs(r_fused - r_pass).max()
    say(f"  max |fused - passes| = {diff:.3e}, max |rhs| = {np.abs(r_pass).max():.3e}")
    assert np.array_equal(r_fused, r_pass)
    assert np.array_equal(uc_fused, uc_pass)
    ctx.close()


def test_mask_change_refreshes_the_plane_flags(monkeypatch):
    """The fused kernel's mask-free flags follow a mask set after a first
    call: the second call, with the sphere, equals the passes with it."""
    N = (70, 18, 40)
    cfg, spec = W.make_config("C2", N)
    spec = dict(spec, immersed=0)
    cfg = dict(cfg, immersed=1)
    coor = W.grid(spec)
    ctx = gpu_case(cfg, spec, coor)
    metrics = {m: ctx.metric(m).host() for m in ("cent", "icsi", "jeta", "kzet")}
    ucat, ucont = W.fields(cfg, spec, metrics, ctx.nvert().host())
    nut = np.zeros_like(ctx.nvert().host())
    rhs0 = W.random_vector(cfg, 0.3, 79)
    _run(ctx, cfg, ucont, ucat, nut, rhs0, monkeypatch, None)   # flags of the empty mask
    ctx.nvert_sphere(*W.SPHERE)
    assert ctx.nvert().host().max() > 0
    r_pass, _ = _run(ctx, cfg, ucont, ucat, nut, rhs0, monkeypatch, "passes")
    r_fused, _ = _run(ctx, cfg, ucont, ucat, nut, rhs0, monkeypatch, None)
    assert np.array_equal(r_fused, r_pass)
    ctx.close()
