"""The five benchmark configurations of the fractional-step hot path as
concrete, seeded workloads (grid, mask, velocity, pressure RHS).

C1  laminar channel, uniform 64^3, Re=180, periodic x/z, no-slip y walls,
    Poiseuille + U[-1,1)*0.05 noise (seed 1001), nvert = 0.
C2  256x128x128 tanh-stretched (beta=2) channel with a sinusoidal bottom wall
    (amp 0.1, fading linearly to a flat top), Re=5000, Smagorinsky LES
    (Cs^2 = 0.01), immersed sphere r=0.25 at (Lx/2, 1, Lz/2); solenoidal
    random-Fourier k^-5/3 field (seed 1002) on a Poiseuille mean.
C3  the C2 generator at 512x256x256 with a random Poisson RHS (seed 1003).
C4  C2 fields with unit-norm random tangent/cotangent (seeds 14001/24001).
C5  768x384x384 turbine wake: tanh-stretched channel (beta=2, flat bottom)
    with a wall-function bottom (bctype -1), a log-law inflow at k = 0 with
    U[-1,1)*0.05 fluctuations (seed 1005, bctype 5), an outlet at k = mz-1
    (bctype 4), Smagorinsky LES and an actuator-disk ring (ACTUATOR_RING).
Box: Lx = 2 pi, Ly = 2, Lz = pi in every configuration.
"""
import math

import numpy as np

from . import channel_grid, loglaw_k, poiseuille, random_scalar, random_vector, turbulent

LX, LY, LZ = 2 * math.pi, 2.0, math.pi

BASE = {
    "C1": dict(N=(64, 64, 64), beta=0.0, amp=0.0, ren=180.0, les=0, immersed=0, field="poiseuille", seed=1001),
    "C2": dict(N=(256, 128, 128), beta=2.0, amp=0.1, ren=5000.0, les=1, immersed=1, field="turbulent", seed=1002),
    "C3": dict(N=(512, 256, 256), beta=2.0, amp=0.1, ren=5000.0, les=1, immersed=1, field="turbulent", seed=1003),
    "C4": dict(N=(256, 128, 128), beta=2.0, amp=0.1, ren=5000.0, les=1, immersed=1, field="turbulent", seed=1002),
    # C1 / C2 with a wall-function bottom wall (bctype -1), as the turbine-wake
    # case has
    "C1w": dict(N=(64, 64, 64), beta=0.0, amp=0.0, ren=5000.0, les=0, immersed=0, field="poiseuille", seed=1001,
                bottom=-1),
    "C2w": dict(N=(256, 128, 128), beta=2.0, amp=0.1, ren=5000.0, les=1, immersed=1, field="turbulent", seed=1002,
                bottom=-1),
    # C1 / C2 with an inlet at k = 0 and an outlet at k = mz-1 (bctypes 5, 4)
    # instead of k periodicity, as the turbine-wake case has
    "C1o": dict(N=(64, 64, 64), beta=0.0, amp=0.0, ren=5000.0, les=0, immersed=0, field="poiseuille", seed=1001,
                kopen=1),
    "C2o": dict(N=(256, 128, 128), beta=2.0, amp=0.1, ren=5000.0, les=1, immersed=1, field="turbulent", seed=1002,
                kopen=1),
    # C1 with a wall-function bottom and a k inlet / outlet (the boundary
    # set of the turbine-wake case, at parity-test size)
    "C1wo": dict(N=(64, 64, 64), beta=0.0, amp=0.0, ren=5000.0, les=0, immersed=0, field="poiseuille", seed=1001,
                 bottom=-1, kopen=1),
    # the turbine-wake case: wall-function bottom, log-law inflow at k = 0,
    # outlet at k = mz-1, LES
    "C5": dict(N=(768, 384, 384), beta=2.0, amp=0.0, ren=5000.0, les=1, immersed=0, field="loglaw", seed=1005,
               bottom=-1, kopen=1),
    # one GPU's share of the turbine-wake case (768 x 384 x 384 cut into 8
    # k slabs): wall-function bottom, log-law inflow at k = 0, outlet at
    # k = mz-1, LES
    "C5s": dict(N=(768, 384, 48), beta=2.0, amp=0.0, ren=5000.0, les=1, immersed=0, field="loglaw", seed=1005,
                bottom=-1, kopen=1),
}
LOGLAW_Y0 = 1e-3         # roughness height of the turbine-wake inflow
CS_LES = 0.01            # Compute_Smagorinsky_Constant_1, les.c:74-77 (Cs = 0.1^2)
SPHERE = (LX / 2, 1.0, LZ / 2, 0.25)


def make_config(name, N=None, dt=1e-3):
    """Return (cfg dict, spec dict) for a configuration, optionally resized to
    N = (Ni, Nj, Nk) cells (tests use reduced sizes of the same generator)."""
    spec = dict(BASE[name])
    if N is not None:
        spec["N"] = tuple(N)
    Ni, Nj, Nk = spec["N"]
    kopen = spec.get("kopen", 0)
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=0, k_periodic=0 if kopen else 1,
               bctype=[7, 7, spec.get("bottom", 1), 1] + ([5, 4] if kopen else [7, 7]), immersed=spec["immersed"],
               les=spec["les"], second_order=0,
               dpdz_set=1, weno=0, ren=spec["ren"], dt=dt, st=1.0, mean_pressure_gradient=0.0)
    spec["name"] = name
    return cfg, spec


def grid(spec):
    Ni, Nj, Nk = spec["N"]
    return channel_grid(Ni + 1, Nj + 1, Nk + 1, LX, LY, LZ, spec["beta"], spec["amp"])


def fields(cfg, spec, metrics, nvert):
    """Velocity (ucat, ucont) for the configuration from host metric arrays
    (dict with cent, icsi, jeta, kzet)."""
    if spec["field"] == "loglaw":
        return loglaw_k(cfg, LY, LOGLAW_Y0, 0.05, spec["seed"], metrics["cent"], nvert, metrics["icsi"],
                        metrics["jeta"], metrics["kzet"])
    if spec["field"] == "poiseuille":
        return poiseuille(cfg, LY, 0.05, spec["seed"], metrics["cent"], nvert, metrics["icsi"], metrics["jeta"],
                          metrics["kzet"])
    return turbulent(cfg, LX, LY, LZ, 24, 16, 0.1, spec["seed"], metrics["cent"], nvert, metrics["icsi"],
                     metrics["jeta"], metrics["kzet"])


def poisson_rhs_random(cfg, seed, nvert, aj):
    return random_scalar(cfg, 1.0, seed, nvert, aj)


def unit_vector(cfg, seed):
    v = random_vector(cfg, 1.0, seed)
    return v / np.linalg.norm(v)


def icosphere(cx, cy, cz, r, level):
    """A triangulated sphere (the body of a CURVIB search): the icosahedron
    with each triangle split in four `level` times, vertices projected onto
    the sphere.  Returns (verts (n_v, 3), elems (n_elmt, 3) int32, outward
    unit normals (n_elmt, 3)); 20 * 4**level elements."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    x = np.array(verts) * r + np.array([cx, cy, cz])
    e = np.array(faces, np.int32)
    n = np.cross(x[e[:, 1]] - x[e[:, 0]], x[e[:, 2]] - x[e[:, 0]])
    n /= np.linalg.norm(n, axis=1)[:, None]
    out = np.einsum("ij,ij->i", n, x[e[:, 0]] - np.array([cx, cy, cz])) < 0
    e[out] = e[out][:, [0, 2, 1]]
    n[out] = -n[out]
    return x, e, n


# the actuator disk of the turbine-wake case: a ring of radius 0.5 about
# (Lx/2, 1, Lz/4) in the x-y plane (normal to the stream, which runs in k),
# thickness 0.1, thrust -1 per unit area along z, Gaussian spreading
ACTUATOR_RING = dict(n=1024, radius=0.5, thickness=0.1, centre=(LX / 2, 1.0, LZ / 4), deltafunc=7, halfwidth=4.0)


def actuator_ring(n=None, radius=None, thickness=None, centre=None, **_):
    """Element centres (3, n), forces (3, n) and areas (n) of the actuator
    ring: n equal arcs, each of area 2 pi r / n * thickness."""
    r = ACTUATOR_RING
    n = n or r["n"]
    radius = radius or r["radius"]
    thickness = thickness or r["thickness"]
    cx, cy, cz = centre or r["centre"]
    th = 2 * np.pi * np.arange(n) / n
    cent = np.stack([cx + radius * np.cos(th), cy + radius * np.sin(th), np.full(n, cz)])
    F = np.stack([np.zeros(n), np.zeros(n), -np.ones(n)])
    dA = np.full(n, 2 * np.pi * radius / n * thickness)
    return cent, F, dA


def loglaw_inflow_plane(cfg, coor, seed=1005, noise=0.05):
    """Inflow speeds of the k = 0 faces (mx * my, j * mx + i) for the
    caller-profile inlet: the log law at the face centre's y (the corners
    InflowFlux averages, bcs.c:386-388) plus U[-1,1) noise."""
    from . import _wl_uniform_array
    mx, my = cfg["mx"], cfg["my"]
    c = coor.reshape(cfg["mz"], my, mx, 3)
    yc = np.zeros((my, mx))
    yc[1:-1, 1:-1] = 0.25 * (c[1, 1:-1, 1:-1, 1] + c[1, :-2, 1:-1, 1] + c[1, 1:-1, :-2, 1] + c[1, :-2, :-2, 1])
    us = 1.0 / ((1.0 / 0.4) * math.log(LY / LOGLAW_Y0))
    w = (1.0 / 0.4) * np.log(np.maximum(yc, LOGLAW_Y0) / LOGLAW_Y0) * us
    return (w + noise * _wl_uniform_array(seed + 1, mx * my).reshape(my, mx)).ravel()
