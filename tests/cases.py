"""Shared set-up of oracle and GPU cases for the parity tests."""
import numpy as np

from oracle_py import Oracle
from vfsgpu import workloads as W

# reduced sizes of the benchmark generators used by the parity tests
SMALL = {
    "C1": (16, 12, 14),
    "C2": (32, 24, 20),
    "C1w": (16, 12, 14),
    "C2w": (32, 24, 20),
    "C1o": (16, 12, 14),
    "C2o": (32, 24, 20),
    "C1wo": (16, 12, 14),
}


def oracle_case(name, N=None, dt=1e-3):
    cfg, spec = W.make_config(name, N or SMALL.get(name), dt=dt)
    coor = W.grid(spec)
    o = Oracle(cfg, coor)
    if spec["immersed"]:
        o.nvert_sphere(*W.SPHERE)
    o.wall_mask()   # IB nodes on wall-function walls (no-op unless bctype -1 and no body)
    ucat, ucont = W.fields(cfg, spec, o.g, o.nvert)
    return cfg, spec, coor, o, ucat, ucont


def gpu_case(cfg, spec, coor):
    import vfsgpu as V
    ctx = V.Ctx(cfg)
    dcoor = V.DeviceArray.from_host(coor)
    ctx.build_metrics(dcoor)
    if spec["immersed"]:
        ctx.nvert_sphere(*W.SPHERE)
    return ctx
