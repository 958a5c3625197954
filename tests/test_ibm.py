"""CURVIB velocity reconstruction: the node list of ibm_search_advanced
(ibm.c:366-497: BoundingSphere 1276-1364, nearestcell1 1366-1511,
InterceptionPoint 1759-2250 with ICP 1659-1757) and
ibm_interpolation_advanced (ibm.c:3294-3698).

CPU tests pin the oracle:
  - the bounding spheres hold their elements, and the sphere of an acute
    element is its circumsphere;
  - d_s is the exact distance from each node to the triangulated body, from
    a closest-point-on-triangle computation of its own (vertex / edge / face
    regions); pmin lies on the element and cs1..3 rebuild it;
  - the interception point lies on the node's normal ray at d_i, and
    cr1..3 rebuild it from the three cell centres;
  - a tilted plane body with a velocity and pressure linear in the distance
    from the plane: noslip gives the field at the node, freeslip the
    tangential part of the interception point and the scaled normal part,
    ustar = sqrt(|shear| / Re), p the value at the interception point;
  - a uniform stream past a sphere on a uniform grid (-inv) comes back
    unchanged at every node, face flux and fluid cell;
  - the list runs from the last cell in k, j, i order to the first;
  - the ray-casting marking of a triangulated body (point_cell_advanced)
    against the exact inside test of a convex polyhedron (every face plane
    has the point on its inner side), for one sphere and for two, with the
    ray jitter equal to the C library's srand / rand.
GPU tests compare the HIP path with the oracle.
"""
import numpy as np
import pytest

from cases import oracle_case
from oracle_py import IbBody, Oracle, lib
from vfsgpu import workloads as W
import vfsgpu as V
import ctypes as C


def sphere_body(level=3, u=None):
    x, e, n = W.icosphere(*W.SPHERE, level=level)
    return x, e, n, u


def closest_on_triangle(p, a, b, c):
    """Closest point to p on triangle abc by its Voronoi regions (vertex,
    edge, face), for many triangles at once (arrays of shape (m, 3))."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = np.einsum("ij,ij->i", ab, ap), np.einsum("ij,ij->i", ac, ap)
    bp = p - b
    d3, d4 = np.einsum("ij,ij->i", ab, bp), np.einsum("ij,ij->i", ac, bp)
    cp = p - c
    d5, d6 = np.einsum("ij,ij->i", ab, cp), np.einsum("ij,ij->i", ac, cp)
    va, vb, vc = d3 * d6 - d5 * d4, d5 * d2 - d1 * d6, d1 * d4 - d3 * d2
    out = np.empty_like(a)
    done = np.zeros(len(a), bool)

    def put(mask, val):
        m = mask & ~done
        out[m] = val[m]
        done[:] |= m

    put((d1 <= 0) & (d2 <= 0), a)
    put((d3 >= 0) & (d4 <= d3), b)
    with np.errstate(divide="ignore", invalid="ignore"):
        put((vc <= 0) & (d1 >= 0) & (d3 <= 0), a + (d1 / (d1 - d3))[:, None] * ab)
        put((d6 >= 0) & (d5 <= d6), c)
        put((vb <= 0) & (d2 >= 0) & (d6 <= 0), a + (d2 / (d2 - d6))[:, None] * ac)
        put((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0),
            b + ((d4 - d3) / ((d4 - d3) + (d5 - d6)))[:, None] * (c - b))
        den = 1.0 / (va + vb + vc)
        put(np.ones(len(a), bool), a + (vb * den)[:, None] * ab + (vc * den)[:, None] * ac)
    return out


def test_bounding_spheres_hold_their_elements():
    x, e, n, _ = sphere_body(2)
    body = IbBody(x, e, n)
    q, r = np.zeros(3), C.c_double(0)
    qa = (C.c_double * 3)()
    for el in range(0, len(e), 7):
        lib().vo_ib_bounding_sphere(C.byref(body.c), C.c_int(el), qa, C.byref(r))
        q = np.array(qa[:])
        d = np.linalg.norm(x[e[el]] - q, axis=1)
        assert np.all(d <= r.value * (1 + 1e-12))
        a, b, c = x[e[el]]
        angles = [np.dot(b - a, c - a), np.dot(a - b, c - b), np.dot(a - c, b - c)]
        if min(angles) > 0:   # acute: the circumcentre, equidistant and in the plane
            assert np.ptp(d) < 1e-12 * r.value
            assert abs(np.dot(q - a, np.cross(b - a, c - a))) < 1e-12 * np.linalg.norm(np.cross(b - a, c - a))


def test_patch_points_follow_the_reference_layout():
    # ibm.c:1785-1797 (i-1 plane), 2058-2070 (j+1 plane), 2194-2206 (k+1 plane)
    ii, jj, kk = C.c_int(), C.c_int(), C.c_int()
    got = []
    for pl in (0, 3, 5):
        row = []
        for q in range(9):
            lib().vo_ib_patch_point(pl, q, 10, 20, 30, C.byref(ii), C.byref(jj), C.byref(kk))
            row.append((ii.value, jj.value, kk.value))
        got.append(row)
    assert got[0] == [(9, 19, 29), (9, 20, 29), (9, 21, 29), (9, 19, 30), (9, 20, 30), (9, 21, 30),
                      (9, 19, 31), (9, 20, 31), (9, 21, 31)]
    assert got[1] == [(9, 21, 29), (10, 21, 29), (11, 21, 29), (9, 21, 30), (10, 21, 30), (11, 21, 30),
                      (9, 21, 31), (10, 21, 31), (11, 21, 31)]
    assert got[2] == [(9, 19, 31), (10, 19, 31), (11, 19, 31), (9, 20, 31), (10, 20, 31), (11, 20, 31),
                      (9, 21, 31), (10, 21, 31), (11, 21, 31)]


@pytest.mark.parametrize("N", [(32, 24, 20), (64, 48, 40)])
def test_search_nearest_point_and_interception(N):
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", N)
    x, e, n, _ = sphere_body(3)
    nodes = o.ib_search(IbBody(x, e, n))
    assert len(nodes) > 20
    mx, my = cfg["mx"], cfg["my"]
    lin = (nodes["k"] * my + nodes["j"]) * mx + nodes["i"]
    assert np.all(np.diff(lin) < 0)                      # AddIBMNode prepends
    nv = o.nvert
    assert sorted(lin) == sorted(np.flatnonzero(np.rint(nv) == 1))
    cent = o.g["cent"].reshape(-1, 3)
    a, b, c = x[e[:, 0]], x[e[:, 1]], x[e[:, 2]]
    for nd, li in zip(nodes, lin):
        p = cent[li]
        cp = closest_on_triangle(np.broadcast_to(p, a.shape), a, b, c)
        dist = np.linalg.norm(cp - p, axis=1)
        assert abs(nd["d_s"] - dist.min()) < 1e-12
        el = nd["cell"]
        assert abs(dist[el] - dist.min()) < 1e-12           # the element holds a nearest point
        assert np.linalg.norm(nd["pmin"] - cp[el]) < 1e-10
        rebuilt = nd["cs1"] * a[el] + nd["cs2"] * b[el] + nd["cs3"] * c[el]
        assert np.linalg.norm(rebuilt - nd["pmin"]) < 1e-12
        # the interception point: on the normal ray at d_i, rebuilt from the cells
        nr = (p - nd["pmin"]) / np.linalg.norm(p - nd["pmin"])
        pint = p + nd["d_i"] * nr
        cells = [((nd[f"k{m}"] * my + nd[f"j{m}"]) * mx + nd[f"i{m}"]) for m in (1, 2, 3)]
        w = [nd["cr1"], nd["cr2"], nd["cr3"]]
        assert abs(sum(w) - 1) < 1e-14
        assert np.linalg.norm(sum(wm * cent[cm] for wm, cm in zip(w, cells)) - pint) < 1e-12
        assert 0 <= nd["imode"] < 8
        for m in (1, 2, 3):
            assert max(abs(nd[f"i{m}"] - nd["i"]), abs(nd[f"j{m}"] - nd["j"]), abs(nd[f"k{m}"] - nd["k"])) == 1


# ---------------------------------------------------------------------------
# a tilted plane body and fields linear in the distance from it
# ---------------------------------------------------------------------------
NORMAL = np.array([0.15, 1.0, -0.1]) / np.linalg.norm([0.15, 1.0, -0.1])
TANGENT = np.cross(NORMAL, [0.0, 0.0, 1.0])
TANGENT /= np.linalg.norm(TANGENT)
PLANE_C = 0.62
U0 = np.array([0.3, -0.2, 0.1])
GAMMA, BETA = 1.7, -0.4
P_A, P_B = 0.5, np.array([0.2, -0.7, 0.3])


def plane_case(N=(24, 20, 18), inviscid=0):
    Ni, Nj, Nk = N
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=0, k_periodic=1,
               bctype=[7, 7, 1, 1, 7, 7], immersed=1, les=0, second_order=0, dpdz_set=1, inviscid=inviscid,
               ren=400.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    coor = V.channel_grid(Ni + 1, Nj + 1, Nk + 1, W.LX, W.LY, W.LZ, 0.0, 0.0)
    o = Oracle(cfg, coor)
    cent = o.g["cent"].reshape(cfg["mz"], cfg["my"], cfg["mx"], 3)
    inner = np.zeros(cent.shape[:3], bool)
    inner[1:-1, 1:-1, 1:-1] = True
    solid = inner & (cent @ NORMAL < PLANE_C)
    near = np.zeros_like(solid)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                near |= np.roll(solid, (dk, dj, di), axis=(0, 1, 2))
    nv = np.where(solid, 3.0, np.where(inner & near, 1.0, 0.0))
    o.nvert[:] = nv.ravel()
    # the plane as 2 x 8 x 8 triangles over a rectangle wider than the box
    t2 = np.cross(NORMAL, TANGENT)
    base = NORMAL * PLANE_C + np.array([W.LX / 2, 0, W.LZ / 2]) - NORMAL * np.dot(NORMAL, [W.LX / 2, 0, W.LZ / 2])
    s = np.linspace(-6, 6, 9)
    verts = np.array([base + a * TANGENT + b * t2 for b in s for a in s])
    elems = []
    for b in range(8):
        for a in range(8):
            v0 = b * 9 + a
            elems += [(v0, v0 + 1, v0 + 10), (v0, v0 + 10, v0 + 9)]
    elems = np.array(elems, np.int32)
    nrm = np.cross(verts[elems[:, 1]] - verts[elems[:, 0]], verts[elems[:, 2]] - verts[elems[:, 0]])
    flip = nrm @ NORMAL < 0
    elems[flip] = elems[flip][:, [0, 2, 1]]
    normals = np.tile(NORMAL, (len(elems), 1))
    uv = np.tile(U0, (len(verts), 1))
    dist = cent @ NORMAL - PLANE_C
    ucat = (U0 + (GAMMA * TANGENT + BETA * NORMAL) * dist[..., None]).ravel()
    p = (P_A + cent @ P_B).ravel()
    return cfg, coor, o, (verts, elems, normals, uv), ucat, p


@pytest.mark.parametrize("inviscid", [0, 1])
def test_plane_body_linear_fields_closed_form(inviscid):
    cfg, coor, o, (verts, elems, normals, uv), ucat, p = plane_case(inviscid=inviscid)
    body = IbBody(verts, elems, normals, uv)
    nodes = o.ib_search(body)
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    cent = o.g["cent"].reshape(-1, 3)
    ucat0, p0 = ucat.copy(), p.copy()
    ucont = o.vec()
    ustar = o.sca()
    cnt, sh, pr = o.ib_interpolate(body, nodes, ucont, ucat, p, ustar)
    assert cnt.sum() == len(nodes)
    checked = nonzero = 0
    for nd in nodes:
        if not (2 <= nd["i"] <= mx - 3 and 2 <= nd["k"] <= mz - 3 and nd["j"] >= 2):
            continue   # patches touching ghost cells
        li = (nd["k"] * my + nd["j"]) * mx + nd["i"]
        x = cent[li]
        sb = x @ NORMAL - PLANE_C
        assert abs(nd["d_s"] - sb) < 1e-12
        sc = sb + nd["d_i"]
        cells = [((nd[f"k{m}"] * my + nd[f"j{m}"]) * mx + nd[f"i{m}"]) for m in (1, 2, 3)]
        got = ucat[3 * li:3 * li + 3]
        assert abs(p[li] - (P_A + (x + nd["d_i"] * NORMAL) @ P_B)) < 1e-12
        if not inviscid and sum(o.nvert[cm] for cm in cells) > 0.1:
            assert np.all(got == 0) and ustar[li] == 0
        elif inviscid:
            want = U0 + GAMMA * sc * TANGENT + BETA * sb * NORMAL
            assert np.abs(got - want).max() < 1e-12
            nonzero += 1
        else:
            want = U0 + (GAMMA * TANGENT + BETA * NORMAL) * sb
            assert np.abs(got - want).max() < 1e-12
            assert abs(ustar[li] - np.sqrt(abs(GAMMA) / cfg["ren"])) < 1e-12
            nonzero += 1
        checked += 1
    assert checked > 50 and nonzero > 10
    # the element means are means over the element's nodes
    lin = (nodes["k"] * my + nodes["j"]) * mx + nodes["i"]
    np.testing.assert_allclose(pr * np.maximum(cnt, 1), np.bincount(nodes["cell"], p[lin], len(elems)), atol=1e-12)
    np.testing.assert_allclose(sh * np.maximum(cnt, 1), np.bincount(nodes["cell"], ustar[lin] ** 2, len(elems)),
                               atol=1e-12)


def test_uniform_stream_survives_the_reconstruction():
    # -inv (no fluid-support guard): every node gets Ua + 0, then the faces
    # and Contra2Cart carry the stream unchanged (a uniform grid, where the
    # cell metric is the mean of its face metrics)
    cfg, spec = W.make_config("C2", (32, 24, 20))
    cfg = dict(cfg, inviscid=1)
    coor = V.channel_grid(33, 25, 21, W.LX, W.LY, W.LZ, 0.0, 0.0)
    o = Oracle(cfg, coor)
    o.nvert_sphere(*W.SPHERE)
    x, e, n, _ = sphere_body(3)
    body = IbBody(x, e, n, np.tile(U0, (len(x), 1)))
    nodes = o.ib_search(body)
    ucat = np.tile(U0, o.P)
    ucont = o.vec()
    for f, nm in enumerate(("icsi", "jeta", "kzet")):
        ucont[f::3] = o.g[nm].reshape(-1, 3) @ U0
    ref = ucont.copy()
    p, ustar = o.sca(), o.sca()
    o.ib_interpolate(body, nodes, ucont, ucat, p, ustar)
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    inner = np.zeros((mz, my, mx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    fluid = (inner & (o.nvert.reshape(mz, my, mx) < 0.1)).ravel()
    assert np.abs(ucat.reshape(-1, 3)[fluid] - U0).max() < 1e-13
    lin = (nodes["k"] * my + nodes["j"]) * mx + nodes["i"]
    assert np.abs(ucat.reshape(-1, 3)[lin] - U0).max() < 1e-15
    assert np.abs(ucont.reshape(-1, 3)[inner.ravel()] - ref.reshape(-1, 3)[inner.ravel()]).max() < 1e-15


# ---------------------------------------------------------------------------
# the ray-casting marking
# ---------------------------------------------------------------------------
SPHERE_B = (W.LX / 2 + 0.6, 0.8, W.LZ / 2 - 0.3, 0.2)


def two_spheres(level=3):
    x1, e1, n1 = W.icosphere(*W.SPHERE, level=level)
    x2, e2, n2 = W.icosphere(*SPHERE_B, level=level)
    return np.vstack([x1, x2]), np.vstack([e1, e2 + len(x1)]), np.vstack([n1, n2])


def convex_inside(cent, x, e, n):
    """Strictly inside the convex polyhedron: every face plane has the point
    on its inner side; None where a point lies within 1e-9 of a plane."""
    d = np.einsum("pk,ek->pe", cent, n) - np.einsum("ek,ek->e", x[e[:, 0]], n)[None, :]
    inside = (d < -1e-9).all(axis=1)
    near = (np.abs(d) <= 1e-9).any(axis=1) & (d <= 1e-9).all(axis=1)
    return inside, near


def test_ray_jitter_is_the_c_library_rand():
    import ctypes.util
    libc = C.CDLL(ctypes.util.find_library("c"))
    for seed in (1, 2, 3, 17, 31):
        libc.srand(seed)
        want = libc.rand() / (2147483647 + 1.0) - 0.5
        assert V.ray_jitter(seed) == want
        assert lib().vo_ib_ray_jitter(seed) == want


@pytest.mark.parametrize("bodies", [1, 2])
def test_ray_casting_marks_the_inside_of_convex_bodies(bodies):
    cfg, spec, coor, o, _, _ = oracle_case("C2", (48, 40, 32))
    o.nvert[:] = 0
    if bodies == 1:
        parts = [W.icosphere(*W.SPHERE, level=3)]
    else:
        parts = [W.icosphere(*W.SPHERE, level=3), W.icosphere(*SPHERE_B, level=3)]
    for q, (x, e, n) in enumerate(parts):
        o.nvert_body(IbBody(x, e, n), reset=(q == 0))
    cent = o.g["cent"].reshape(-1, 3)
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    inner = np.zeros((mz, my, mx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.ravel()
    want = np.zeros(o.P, bool)
    for x, e, n in parts:
        ins, near = convex_inside(cent, x, e, n)
        assert not near[inner].any()
        want |= ins
    got = np.rint(o.nvert)
    assert np.array_equal(got[inner] == 3, want[inner])
    # the dilation: 1 exactly where a solid cell is in the 3x3x3 box
    sol = (got == 3).reshape(mz, my, mx)
    near = np.zeros_like(sol)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                sh = np.zeros_like(sol)
                sh[max(dk, 0):mz + min(dk, 0), max(dj, 0):my + min(dj, 0), max(di, 0):mx + min(di, 0)] = \
                    sol[max(-dk, 0):mz - max(dk, 0), max(-dj, 0):my - max(dj, 0), max(-di, 0):mx - max(di, 0)]
                near |= sh
    ib = (near & ~sol).ravel() & inner
    assert np.array_equal(got[inner] == 1, ib[inner])
    assert (got == 3).sum() > 20


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def gpu_surface(x, e, n, u=None):
    return V.IbSurface(x, e, n, u)


def nodes_equal(a, b):
    for f in a.dtype.names:
        assert np.array_equal(a[f], b[f]), f


def run_pair(cfg, coor, o, body_arrays, ucat, ucont, p, nvert_gpu=None):
    x, e, n, u = body_arrays
    body = IbBody(x, e, n, u)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(o.nvert))
    surf = gpu_surface(x, e, n, u)
    nodes_o = o.ib_search(body)
    dn, cnt_n = ctx.ib_search(surf)
    assert cnt_n == len(nodes_o)
    nodes_g = dn.host()[:cnt_n * V.IB_NODE_DTYPE.itemsize].view(V.IB_NODE_DTYPE)
    nodes_equal(nodes_g, nodes_o)
    ustar = np.zeros(o.P)
    d_ucat, d_ucont, d_p, d_ustar = (V.DeviceArray.from_host(a) for a in (ucat, ucont, p, ustar))
    st = ctx.ib_interpolate(surf, dn, cnt_n, d_ucont, d_ucat, d_p, d_ustar)
    ucat_o, ucont_o, p_o = ucat.copy(), ucont.copy(), p.copy()
    cnt, sh, pr = o.ib_interpolate(body, nodes_o, ucont_o, ucat_o, p_o, ustar)
    assert np.array_equal(d_ucat.host(), ucat_o)
    assert np.array_equal(d_ucont.host(), ucont_o)
    assert np.array_equal(d_p.host(), p_o)
    assert np.array_equal(d_ustar.host(), ustar)
    assert np.array_equal(st[0].host(), cnt)
    assert np.array_equal(st[1].host(), sh)
    assert np.array_equal(st[2].host(), pr)
    return len(nodes_o)


@pytest.mark.gpu
@pytest.mark.parametrize("name,inviscid", [("C2", 0), ("C2", 1), ("C2w", 0)])
def test_gpu_ib_reconstruction_sphere(name, inviscid):
    cfg, spec, coor, o0, ucat, ucont = oracle_case(name)
    cfg = dict(cfg, inviscid=inviscid)
    o = Oracle(cfg, coor)
    o.nvert[:] = o0.nvert
    x, e, n, _ = sphere_body(3)
    u = 0.05 * np.sin(np.arange(3 * len(x))).reshape(-1, 3)
    p = np.sin(np.arange(o.P) * 0.37)
    n_nodes = run_pair(cfg, coor, o, (x, e, n, u), ucat, ucont, p)
    assert n_nodes > 20


@pytest.mark.gpu
def test_gpu_ib_reconstruction_full_size():
    # 256 x 128 x 128 with the sphere as 5120 triangles
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", W.BASE["C2"]["N"])
    x, e, n, _ = sphere_body(4)
    p = np.cos(np.arange(o.P) * 0.013)
    n_nodes = run_pair(cfg, coor, o, (x, e, n, None), ucat, ucont, p)
    print("IB nodes", n_nodes, flush=True)
    assert n_nodes > 1000


@pytest.mark.gpu
@pytest.mark.parametrize("inviscid", [0, 1])
def test_gpu_ib_reconstruction_plane(inviscid):
    cfg, coor, o, (verts, elems, normals, uv), ucat, p = plane_case(inviscid=inviscid)
    ucont = np.sin(np.arange(3 * o.P) * 0.11)
    assert run_pair(cfg, coor, o, (verts, elems, normals, uv), ucat, ucont, p) > 100


@pytest.mark.gpu
@pytest.mark.parametrize("bodies", [1, 2])
def test_gpu_nvert_body(bodies):
    cfg, spec, coor, o, _, _ = oracle_case("C2", (48, 40, 32))
    o.nvert[:] = 0
    parts = [W.icosphere(*W.SPHERE, level=3)] + ([W.icosphere(*SPHERE_B, level=3)] if bodies == 2 else [])
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    for q, (x, e, n) in enumerate(parts):
        o.nvert_body(IbBody(x, e, n), reset=(q == 0))
        got = ctx.nvert_body(V.IbSurface(x, e, n), reset=(q == 0))
    assert np.array_equal(got.host(), o.nvert)
    assert np.array_equal(ctx.nvert().host(), o.nvert)


@pytest.mark.gpu
def test_gpu_nvert_body_full_size_matches_search_chain():
    # the triangulated sphere at 256 x 128 x 128: marking, node list and
    # reconstruction all against the oracle
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", W.BASE["C2"]["N"])
    x, e, n = W.icosphere(*W.SPHERE, level=4)
    o.nvert[:] = 0
    o.nvert_body(IbBody(x, e, n))
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    got = ctx.nvert_body(V.IbSurface(x, e, n))
    assert np.array_equal(got.host(), o.nvert)
    n_nodes = run_pair(cfg, coor, o, (x, e, n, None), ucat, ucont, np.cos(np.arange(o.P) * 0.013))
    print("IB nodes of the triangulated sphere", n_nodes, flush=True)


@pytest.mark.gpu
def test_gpu_ib_interpolate_skips_entries_outside_the_grid():
    # one record with an element and an interpolation cell out of range is
    # skipped; the others give what the list without it gives
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    x, e, n, _ = sphere_body(3)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(o.nvert))
    surf = gpu_surface(x, e, n)
    dn, nn = ctx.ib_search(surf)
    nodes = dn.host()[:nn * V.IB_NODE_DTYPE.itemsize].view(V.IB_NODE_DTYPE).copy()
    bad = nodes.copy()
    bad[3]["cell"] = -5
    bad[3]["i1"] = 10 ** 6
    keep = np.delete(nodes, 3)
    outs = []
    for lst in (bad, keep):
        d_ucat, d_ucont, d_p, d_ustar = (V.DeviceArray.from_host(a) for a in (ucat, ucont, np.zeros(o.P), np.zeros(o.P)))
        dl = V.DeviceArray.from_host(lst.view(np.uint8))
        ctx.ib_interpolate(surf, dl, len(lst), d_ucont, d_ucat, d_p, d_ustar, stats=False)
        outs.append((d_ucat.host(), d_ucont.host(), d_p.host()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# the IB wall model (wall_function_roughness_a) and moving bodies
# ---------------------------------------------------------------------------
def test_rough_wall_law_regimes():
    # sediment_transport.c:18-68: viscous sublayer u = utau y+ (y+ <= 11.53);
    # smooth log law u+ = ln(y+)/0.41 + 5.5 (k+ <= 2.25); fully rough
    # u+ = ln(y/ks)/0.41 + 8.5 (k+ >= 90); the Newton inverse gives utau back
    hu, ht = lib().vo_hydset_u, lib().vo_hydset_utau
    nu = 1e-4
    ut = 0.02
    y = 5 * nu / ut
    assert abs(hu(nu, y, ut, 1e-9) - ut * ut * y / nu) < 1e-15
    y = 1e3 * nu / ut
    assert abs(hu(nu, y, ut, 1e-9) - ut * (np.log(ut * y / nu) / 0.41 + 5.5)) < 1e-12
    ks = 200 * nu / ut
    assert abs(hu(nu, y, ut, ks) - ut * (np.log(y / ks) / 0.41 + 8.5)) < 1e-12
    for (u, yy, k) in [(0.5, 0.01, 1e-4), (0.05, 0.002, 1e-4), (1.2, 0.05, 1e-3), (0.3, 0.004, 1e-6)]:
        t = ht(nu, u, yy, 0.01, k)
        assert abs(hu(nu, yy, t, k) - u) < 1e-6 * u


@pytest.mark.parametrize("moving", [False, True])
def test_plane_body_wall_model_and_moving(moving):
    cfg, coor, o, (verts, elems, normals, uv), ucat, p = plane_case()
    body = IbBody(verts, elems, normals, uv)
    nodes = o.ib_search(body)
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    cent = o.g["cent"].reshape(-1, 3)
    ucont, ustar = o.vec(), o.sca()
    u = ucat.copy()
    o.ib_interpolate(body, nodes, ucont, u, p.copy(), ustar, moving=moving, wallfunction=not moving, ks=1e-4)
    nu = 1.0 / cfg["ren"]
    hu, ht = lib().vo_hydset_u, lib().vo_hydset_utau
    checked = 0
    for nd in nodes:
        if not (2 <= nd["i"] <= mx - 3 and 2 <= nd["k"] <= mz - 3 and nd["j"] >= 2):
            continue
        li = (nd["k"] * my + nd["j"]) * mx + nd["i"]
        cells = [((nd[f"k{m}"] * my + nd[f"j{m}"]) * mx + nd[f"i{m}"]) for m in (1, 2, 3)]
        sb = nd["d_s"]
        sc = sb + nd["d_i"]
        got = u[3 * li:3 * li + 3]
        if not moving and sum(o.nvert[cm] for cm in cells) > 0.1:
            assert np.all(got == 0)
            continue
        if moving:   # no guard: the linear field back at every node
            want = U0 + (GAMMA * TANGENT + BETA * NORMAL) * sb
        else:        # the law's speed at s_b for the u_tau that gives gamma s_c at s_c
            t = ht(nu, abs(GAMMA) * sc, sc, 0.01, 1e-4)
            assert abs(ustar[li] - t) < 1e-15
            want = U0 + np.sign(GAMMA) * hu(nu, sb, t, 1e-4) * TANGENT + BETA * sb * NORMAL
        assert np.abs(got - want).max() < 1e-11
        checked += 1
    assert checked > 50


def test_rough_loglaw_closed_form():
    """u = u_tau (ln(y/ks)/0.41 + 8.5) is linear in u_tau, so the Newton
    solve must land on u / (ln(y/ks)/0.41 + 8.5)."""
    ru, rt = lib().vo_roughlog_u, lib().vo_roughlog_utau
    for (u, y, k) in [(0.5, 0.01, 1e-3), (0.05, 0.002, 1e-4), (1.2, 0.05, 2e-3), (3e-4, 0.004, 1e-6)]:
        a = np.log(y / k) / 0.41 + 8.5
        assert abs(ru(y, 0.7, k) - 0.7 * a) < 1e-14 * a
        assert abs(rt(u, y, 0.01, k) - u / a) < 1e-12 * u / a


def test_plane_body_rough_loglaw():
    """The rough log law at the nodes of the plane body: the tangential speed
    gamma s_c at s_c becomes A(s_b)/A(s_c) of it at s_b, A(y) = ln(y/ks)/0.41
    + 8.5, the normal part is dropped, and u_tau = gamma s_c / A(s_c)."""
    cfg, coor, o, (verts, elems, normals, uv), ucat, p = plane_case()
    body = IbBody(verts, elems, normals, uv)
    nodes = o.ib_search(body)
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    ucont, ustar = o.vec(), o.sca()
    u = ucat.copy()
    R = 2e-4
    o.ib_interpolate(body, nodes, ucont, u, p.copy(), ustar, wallfunction=True, ks=1e-4, roughness=R)
    A = lambda y: np.log(y / R) / 0.41 + 8.5
    checked = 0
    for nd in nodes:
        if not (2 <= nd["i"] <= mx - 3 and 2 <= nd["k"] <= mz - 3 and nd["j"] >= 2):
            continue
        li = (nd["k"] * my + nd["j"]) * mx + nd["i"]
        cells = [((nd[f"k{m}"] * my + nd[f"j{m}"]) * mx + nd[f"i{m}"]) for m in (1, 2, 3)]
        sb = nd["d_s"]
        sc = sb + nd["d_i"]
        got = u[3 * li:3 * li + 3]
        if sum(o.nvert[cm] for cm in cells) > 0.1:
            assert np.all(got == 0)
            continue
        assert abs(ustar[li] - abs(GAMMA) * sc / A(sc)) < 1e-12
        want = U0 + GAMMA * sc * A(sb) / A(sc) * TANGENT
        assert np.abs(got - want).max() < 1e-11
        checked += 1
    assert checked > 50


@pytest.mark.gpu
@pytest.mark.parametrize("moving,wallfunction", [(True, False), (False, True), (True, True), (False, "loglaw")])
def test_gpu_ib_options(moving, wallfunction):
    cfg, coor, o, (verts, elems, normals, uv), ucat, p = plane_case()
    body = IbBody(verts, elems, normals, uv)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(o.nvert))
    surf = gpu_surface(verts, elems, normals, uv)
    nodes = o.ib_search(body)
    dn, nn = ctx.ib_search(surf)
    ucont = np.sin(np.arange(3 * o.P) * 0.11)
    d_ucat, d_ucont, d_p, d_us = (V.DeviceArray.from_host(a) for a in (ucat, ucont, p, np.zeros(o.P)))
    rough = 2e-4 if wallfunction == "loglaw" else None
    # ks is read by the smooth-wall law only: the rough log law runs with none
    ks = 0.0 if wallfunction == "loglaw" else 1e-4
    st = ctx.ib_interpolate(surf, dn, nn, d_ucont, d_ucat, d_p, d_us, moving=moving, wallfunction=wallfunction,
                            ks=ks, roughness=rough)
    uc, u, pp, us = ucont.copy(), ucat.copy(), p.copy(), np.zeros(o.P)
    cnt, sh, pr = o.ib_interpolate(body, nodes, uc, u, pp, us, moving=moving, wallfunction=wallfunction, ks=ks,
                                   roughness=rough)
    # the wall law's log / exp / sin: the device's and the C library's
    tol = 1e-12 if wallfunction else 0.0
    for g, h in ((d_ucat.host(), u), (d_ucont.host(), uc), (d_p.host(), pp), (d_us.host(), us)):
        assert np.abs(g - h).max() <= tol * max(1.0, np.abs(h).max())
    assert np.array_equal(st[0].host(), cnt)
    assert np.abs(st[1].host() - sh).max() <= tol * max(1.0, np.abs(sh).max())


# ---------------------------------------------------------------------------
# thin bodies (-thin)
# ---------------------------------------------------------------------------
def thin_plate(n=12):
    """A double-sided plate z = 0.4 + 0.35 x over the whole domain (each
    triangle twice with opposite normals, so rays through it count two
    crossings).  Its box spans the domain, so the 40^3 element bins are
    about as large as the cells, which the bin neighbourhood of
    point_cell_thin needs."""
    xs, ys = np.linspace(-0.2, W.LX + 0.2, n + 1), np.linspace(0.2, 1.8, n + 1)
    verts = np.array([(x, y, 0.4 + 0.35 * x) for y in ys for x in xs])
    el = []
    for b in range(n):
        for a in range(n):
            v = b * (n + 1) + a
            el += [(v, v + 1, v + n + 2), (v, v + n + 2, v + n + 1)]
    el = np.array(el, np.int32)
    nrm = np.array([-0.35, 0.0, 1.0]) / np.hypot(0.35, 1.0)
    elems = np.vstack([el, el[:, [0, 2, 1]]])
    normals = np.vstack([np.tile(nrm, (len(el), 1)), np.tile(-nrm, (len(el), 1))])
    # orient each copy so its normal follows the right-hand rule
    x = verts
    c = np.cross(x[elems[:, 1]] - x[elems[:, 0]], x[elems[:, 2]] - x[elems[:, 0]])
    flip = np.einsum("ij,ij->i", c, normals) < 0
    elems[flip] = elems[flip][:, [0, 2, 1]]
    return verts, elems, normals


def plate_case():
    N = (24, 20, 18)
    cfg = dict(mx=N[0] + 2, my=N[1] + 2, mz=N[2] + 2, i_periodic=1, j_periodic=0, k_periodic=1,
               bctype=[7, 7, 1, 1, 7, 7], immersed=1, les=0, second_order=0, dpdz_set=1, ren=400.0, dt=1e-3, st=1.0,
               mean_pressure_gradient=0.0)
    coor = V.channel_grid(N[0] + 1, N[1] + 1, N[2] + 1, W.LX, W.LY, W.LZ, 0.0, 0.0)
    return cfg, coor, thin_plate()


def test_thin_plate_marks_only_cells_at_the_plate():
    # every IB node of the -thin marking lies in the 2x2x2 box of a cell one of
    # whose four segments (+i, +j, +k, +ijk) crosses the plate inside it
    cfg, coor, (verts, elems, normals) = plate_case()
    o = Oracle(cfg, coor)
    o.nvert_body(IbBody(verts, elems, normals), thin=True)
    mz, my, mx = cfg["mz"], cfg["my"], cfg["mx"]
    nv = np.rint(o.nvert.reshape(mz, my, mx))
    assert not (nv == 3).any()                  # double-sided: nothing inside
    cent = o.g["cent"].reshape(mz, my, mx, 3)
    f = lambda p: p[..., 2] - 0.4 - 0.35 * p[..., 0]
    cross = np.zeros((mz, my, mx), bool)
    for di, dj, dk in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)):
        a = cent[1:-1, 1:-1, 1:-1]
        b = cent[1 + dk:mz - 1 + dk, 1 + dj:my - 1 + dj, 1 + di:mx - 1 + di]
        fa, fb = f(a), f(b)
        with np.errstate(divide="ignore", invalid="ignore"):
            t = fa / (fa - fb)
            pt = a + t[..., None] * (b - a)
        hit = (fa * fb < 0) & (pt[..., 1] > 0.2) & (pt[..., 1] < 1.8)
        cross[1:-1, 1:-1, 1:-1] |= hit
    allowed = np.zeros_like(cross)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                allowed[dk:, dj:, di:] |= cross[:mz - dk, :my - dj, :mx - di]
    marked = nv == 1
    assert marked.sum() > 100
    assert not (marked & ~allowed).any()
    # without -thin a double-sided plate leaves the mask empty
    o2 = Oracle(cfg, coor)
    o2.nvert_body(IbBody(verts, elems, normals), thin=False)
    assert np.all(o2.nvert == 0)


@pytest.mark.gpu
def test_gpu_nvert_thin_body():
    cfg, coor, (verts, elems, normals) = plate_case()
    o = Oracle(cfg, coor)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    x, e, n = W.icosphere(W.LX / 2, 1.0, 2.2, 0.3, 3)
    for q, (vv, ee, nn, thin) in enumerate([(x, e, n, False), (verts, elems, normals, True)]):
        o.nvert_body(IbBody(vv, ee, nn), reset=(q == 0), thin=thin)
        got = ctx.nvert_body(V.IbSurface(vv, ee, nn), reset=(q == 0), thin=thin)
    assert np.array_equal(got.host(), o.nvert)
    assert (np.rint(o.nvert) == 1).sum() > 50 and (np.rint(o.nvert) == 3).sum() > 0


# ---------------------------------------------------------------------------
# the averaging statistics (ibm.c:3481-3521)
# ---------------------------------------------------------------------------
TAU = dict(uu=0.3, vv=0.2, ww=0.1, uv=0.05, vw=-0.02, uw=0.07)


def averaged_sums(o, N, linear):
    """Running sums over N steps of a mean field (U0, or U0 plus the linear
    shear of the plane case) with the constant fluctuation tensor TAU."""
    cent = o.g["cent"].reshape(-1, 3)
    dist = cent @ NORMAL - PLANE_C
    U = U0 + ((GAMMA * TANGENT + BETA * NORMAL) * dist[:, None] if linear else 0 * dist[:, None])
    u2 = np.stack([U[:, 0] ** 2 + TAU["uu"], U[:, 1] ** 2 + TAU["vv"], U[:, 2] ** 2 + TAU["ww"]], 1)
    u1 = np.stack([U[:, 0] * U[:, 1] + TAU["uv"], U[:, 1] * U[:, 2] + TAU["vw"], U[:, 2] * U[:, 0] + TAU["uw"]], 1)
    return (N * U).ravel(), (N * u1).ravel(), (N * u2).ravel()


def test_averages_closed_form():
    # a uniform mean stream carried by the body: ustar_avg = 0, and on a
    # uniform grid the contravariant stresses of csi-eta, eta-zet, csi-zet are
    # the Cartesian uv, vw, uw
    cfg, coor, o, (verts, elems, normals, uv), _, _ = plane_case()
    body = IbBody(verts, elems, normals, uv)
    nodes = o.ib_search(body)
    N = 40.0
    usum, u1sum, u2sum = averaged_sums(o, N, linear=False)
    ms, r1, r2, r3 = o.ib_averages(body, nodes, usum, u1sum, u2sum, N)
    hit = np.bincount(nodes["cell"], minlength=len(elems)) > 0
    assert hit.sum() > 10
    assert np.abs(ms[hit]).max() < 1e-15       # roundoff of U0 - Ua
    assert np.abs(r1[hit] - TAU["uv"]).max() < 1e-12
    assert np.abs(r2[hit] - TAU["vw"]).max() < 1e-12
    assert np.abs(r3[hit] - TAU["uw"]).max() < 1e-12
    # the linear shear: noslip on the mean gives ustar_avg^2 = |gamma| / Re
    # (elements whose nodes' patches stay off the ghost cells)
    usum, u1sum, u2sum = averaged_sums(o, N, linear=True)
    ms, r1, r2, r3 = o.ib_averages(body, nodes, usum, u1sum, u2sum, N)
    mx, mz = cfg["mx"], cfg["mz"]
    edge = (nodes["i"] < 2) | (nodes["i"] > mx - 3) | (nodes["k"] < 2) | (nodes["k"] > mz - 3) | (nodes["j"] < 2)
    clean = hit & (np.bincount(nodes["cell"][edge], minlength=len(elems)) == 0)
    assert clean.sum() > 5
    assert np.abs(ms[clean] - abs(GAMMA) / cfg["ren"]).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("wallfunction", [False, True])
def test_gpu_ib_averages(wallfunction):
    cfg, coor, o, (verts, elems, normals, uv), _, _ = plane_case()
    body = IbBody(verts, elems, normals, uv)
    ctx = V.Ctx(cfg)
    ctx.build_metrics(V.DeviceArray.from_host(coor))
    ctx.set_nvert(V.DeviceArray.from_host(o.nvert))
    surf = gpu_surface(verts, elems, normals, uv)
    nodes = o.ib_search(body)
    dn, nn = ctx.ib_search(surf)
    N = 40.0
    sums = averaged_sums(o, N, linear=True)
    got = ctx.ib_averages(surf, dn, nn, *[V.DeviceArray.from_host(a) for a in sums], N, wallfunction=wallfunction)
    want = o.ib_averages(body, nodes, *sums, N, wallfunction=wallfunction)
    tol = 1e-12 if wallfunction else 0.0
    for g, h in zip(got, want):
        assert np.abs(g.host() - h).max() <= tol * max(1.0, np.abs(h).max())
