"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle on small seeded inputs, with solid, IB-node and fluid cells
in the C2 case).

CPU: the oracle reproduces every fixture output bit for bit (regression
anchor for the oracle).  GPU: the library, fed the fixture's inputs,
reproduces the outputs within 1e-12 relative.
"""
import glob
import os

import numpy as np
import pytest

from conftest import rel_err
from oracle_py import Oracle
from vfsgpu import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
sys_path_golden = os.path.join(HERE, "golden")


def load(path):
    name = os.path.basename(path).split("_")[0].upper()
    N = tuple(int(x) for x in os.path.basename(path).split("_")[1].split(".")[0].split("x"))
    cfg, spec = W.make_config(name, N)
    z = np.load(path)
    return name, cfg, spec, {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) >= 2


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    import sys
    sys.path.insert(0, sys_path_golden)
    from make_golden import compute
    name, cfg, spec, z = load(path)
    assert np.array_equal(W.grid(spec), z["coor"])
    o = Oracle(cfg, z["coor"])
    o.nvert[:] = z["nvert"]
    out = compute(o, cfg, z["ucont"], z["ucat"])
    for k, v in out.items():
        assert np.array_equal(v, z[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_fixture(path):
    import vfsgpu as V
    name, cfg, spec, z = load(path)
    print(f"[golden] {os.path.basename(path)}", flush=True)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(z["coor"]))
    ctx.set_nvert(V.DeviceArray.from_host(z["nvert"]))
    uc, u = V.DeviceArray.from_host(z["ucont"]), V.DeviceArray.from_host(z["ucat"])
    ctx.contra2cart(uc, u)
    assert rel_err(u.host(), z["c2c_ucat"]) < 1e-12
    assert rel_err(uc.host(), z["c2c_ucont"]) < 1e-12
    nut = ctx.sca()
    if cfg["les"]:
        ctx.eddy_viscosity(u, W.CS_LES, nut)
        assert rel_err(nut.host(), z["nu_t"]) < 1e-12
    rhs = ctx.vec()
    ctx.formfunction2(uc, u, nut, rhs, 1.0)
    assert rel_err(rhs.host(), z["ff2_rhs"]) < 1e-12
    dp = ctx.vec()
    ctx.pressure_gradient(V.DeviceArray.from_host(z["p_in"]), dp)
    assert rel_err(dp.host(), z["dp"]) < 1e-12
    ctx.poisson_setup()
    u2 = V.DeviceArray.from_host(z["c2c_ucont"])
    ctx.volume_flux(u2)
    b = ctx.sca()
    ctx.poisson_rhs(u2, b)
    assert rel_err(b.host(), z["poisson_b"]) < 1e-12
    u3 = V.DeviceArray.from_host(z["c2c_ucont"])
    ctx.projection(V.DeviceArray.from_host(z["phi_in"]), u3)
    assert rel_err(u3.host(), z["projection_ucont"]) < 1e-12
    m = ctx.divergence(V.DeviceArray.from_host(z["c2c_ucont"]))
    assert rel_err(m, float(z["maxdiv"][0])) < 1e-12
    ctx.close()
