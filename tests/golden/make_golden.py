"""Writes the golden fixtures of the hot path: small seeded inputs and the
CPU oracle's outputs for them, one .npz per case.

The reference ships no golden vectors for this path and cannot be built
here (PETSc 3.1 / HYPRE / MPI are absent), so these fixtures come from the
oracle, whose functions are pinned by tests/test_oracle.py.  They freeze
the oracle's results (tests/test_golden.py recomputes them bit for bit) and
give the GPU a fixed-input parity case that does not depend on the workload
generators.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "vfs-geophysics-3.1_amd"))

from cases import oracle_case  # noqa: E402
from vfsgpu import workloads as W  # noqa: E402

CASES = {"C1": (8, 8, 8), "C2": (12, 10, 8)}
# at 12x10x8 cells the C2 sphere (r = 0.25) holds no cell centre: the fixture
# uses r = 0.8 so that solid, IB-node and fluid cells all occur
SPHERE_R = {"C2": 0.8}


def compute(o, cfg, ucont, ucat):
    """Run the oracle routines of the path on (ucont, ucat); return outputs."""
    out = {}
    uc, u = ucont.copy(), ucat.copy()
    o.contra2cart(uc, u)
    out["c2c_ucont"], out["c2c_ucat"] = uc.copy(), u.copy()
    nut = o.sca()
    if cfg["les"]:
        o.eddy_viscosity(u, W.CS_LES, nut)
    out["nu_t"] = nut.copy()
    rhs = o.vec()
    o.formfunction2(uc, u, nut, rhs, 1.0)
    out["ff2_rhs"] = rhs
    p = W.random_scalar(cfg, 1.0, 7001)
    dp = o.vec()
    o.pressure_gradient(p.copy(), dp)
    out["dp"] = dp
    o.setup_gid()
    b = o.sca()
    u2 = uc.copy()
    o.volume_flux(u2)
    o.poisson_rhs(u2, b)
    out["poisson_b"] = b
    phi = W.random_scalar(cfg, 0.01, 7002)
    ph = phi.reshape(o.mz, o.my, o.mx)       # periodic ghosts in i and k, as a solve leaves them
    ph[:, :, 0] = ph[:, :, -2]; ph[:, :, -1] = ph[:, :, 1]; ph[0] = ph[-2]; ph[-1] = ph[1]
    u3 = uc.copy()
    o.projection(phi.copy(), u3)
    out["projection_ucont"] = u3
    m, div = o.divergence(uc)
    out["maxdiv"] = np.array([m])
    out["p_in"], out["phi_in"] = p, phi
    return out


def main():
    for name, N in CASES.items():
        cfg, spec, coor, o, ucat, ucont = oracle_case(name, N)
        if name in SPHERE_R:
            cx, cy, cz, _ = W.SPHERE
            o.nvert_sphere(cx, cy, cz, SPHERE_R[name])
            ucat, ucont = W.fields(cfg, spec, o.g, o.nvert)
        out = compute(o, cfg, ucont, ucat)
        path = os.path.join(HERE, f"{name.lower()}_{N[0]}x{N[1]}x{N[2]}.npz")
        np.savez_compressed(path, coor=coor, nvert=o.nvert, ucont=ucont, ucat=ucat, **out)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
