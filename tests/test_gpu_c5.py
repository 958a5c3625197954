"""The turbine-wake case C5 at full size (768x384x384, 113 M cells) on one
GPU: wall-function bottom, log-law inflow with fluctuations at k = 0, the
outlet at k = mz-1, Smagorinsky LES, the actuator ring spread with the
Gaussian delta function and carried in every Runge-Kutta stage, and the
mixed-precision Poisson solve, for 20 steps of Flow_Solver's explicit path.

The reference cannot run this case here, so the steps are checked through
properties that hold at any size: every Poisson solve reaches its residual,
the outflow equals the inflow after FormBCS's correction, the projection
leaves the divergence at the level of the solve, the fields stay finite, and
the actuator force decelerates the stream through the disk.
"""
import sys

import numpy as np
import pytest

from vfsgpu import workloads as W
from vfsgpu.step import FlowSolver
import vfsgpu as V

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_c5_twenty_steps():
    N = W.BASE["C5"]["N"]
    say(f"[c5] C5 {N}")
    cfg, spec = W.make_config("C5", N)
    ctx = V.Ctx(cfg)
    coor = W.grid(spec)
    dcoor = V.DeviceArray.from_host(coor)
    ctx.build_metrics(dcoor)
    nvert = ctx.nvert().host()
    metrics = {m: ctx.metric(m).host() for m in ("cent", "icsi", "jeta", "kzet")}
    ucat, ucont = W.fields(cfg, spec, metrics, nvert)
    del metrics
    plane = V.DeviceArray.from_host(W.loglaw_inflow_plane(cfg, coor))
    del coor
    n = ctx.poisson_setup()
    _, _, inlet_area = ctx.inlet_area()
    say(f"[c5] {n} unknowns, inlet area {inlet_area:.4f}")
    ring = W.ACTUATOR_RING
    cent, F, dA = W.actuator_ring()
    dcent = V.DeviceArray.from_host(cent.reshape(-1))
    box, _ = ctx.rotor_elements(dcoor, dcent)
    f_eul = ctx.vec()
    ctx.calc_f_eul(dcoor, dcent, V.DeviceArray.from_host(F.reshape(-1)), V.DeviceArray.from_host(dA), box, f_eul,
                   deltafunc=ring["deltafunc"], halfwidth=ring["halfwidth"])
    fz = f_eul.host().reshape(cfg["mz"], cfg["my"], cfg["mx"], 3)[..., 2]
    assert fz.min() < 0 and fz.max() <= 0   # thrust against the stream only
    ctx.set_body_force(f_eul)
    fs = FlowSolver(ctx, V.DeviceArray.from_host(ucont), V.DeviceArray.from_host(ucat), coor=dcoor,
                    inflow=V.InflowOptions(0, 0, -1.0, inlet_area), uin_plane=plane, cs=W.CS_LES, precision=1,
                    rtol=1e-8, add_body_force=True)
    del ucat, ucont
    kd = int(np.argmin(np.abs(np.linspace(0, W.LZ, cfg["mz"] - 1) - ring["centre"][2])))
    for step in range(20):
        out = fs.step()
        say(f"[c5] step {step}: {out}")
        assert np.isfinite(out["maxdiv"]) and out["relres"] <= 1e-8
        # measured: 30-34 iterations on the first three steps, 18-24 after
        # (semi-coarsened hierarchy with the over-corrected W-cycle)
        assert out["iters"] <= (45 if step < 3 else 32)
        assert abs(out["flux_out_corrected"] - out["flux_in"]) <= 1e-12 * abs(out["flux_in"])
        # measured: 1e-5 and 2e-6 on the first two steps (the seeded field is
        # not solenoidal), then 1e-7 to 1e-9
        assert out["maxdiv"] <= (1e-4 if step < 2 else 1e-6)
    uc = fs.uc.host().reshape(cfg["mz"], cfg["my"], cfg["mx"], 3)
    assert np.isfinite(uc).all()
    # the disk decelerates the stream: the k flux through the disk area drops
    # just behind it relative to just in front of it
    mask = np.abs(fz).sum(axis=0) > 0
    ahead, behind = uc[kd - 8, ..., 2][mask].sum(), uc[kd + 8, ..., 2][mask].sum()
    say(f"[c5] k flux through the disk's footprint: {ahead:.6f} ahead, {behind:.6f} behind")
    assert behind < ahead
    ctx.close()
