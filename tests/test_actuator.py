"""Actuator force spreading: the element search of the rotor pre-processing
(rotor_model.c:2590-2737), Calc_F_eul (rotor_model.c:4792-5137) with its
discrete delta functions (6042-6200), and the body force in the implicit
residual (momentum.c:2288).

CPU tests pin the oracle:
  - the delta functions: partition of unity and zero first moment on the
    integer lattice (the properties the kernels are built to have), and
    their closed forms at known points;
  - on a uniform Cartesian grid one element's spread force integrates back
    to F dA in each direction (sum of f_eul . width), and its centroid over
    the face centres is the element centre;
  - rigidly rotating the curvilinear grid, the element and its force leaves
    the contravariant f_eul unchanged;
  - the element search equals a brute-force first minimum over the nodes on
    grids the +-21 window covers;
  - the nvert rule: zero force where the 3x3x3 mask sum exceeds 2.9, the
    same force elsewhere.
GPU tests compare the HIP path with the oracle.
"""
import numpy as np
import pytest

from cases import gpu_case, oracle_case
from conftest import rel_err
from oracle_py import Oracle, dfunc
from vfsgpu import workloads as W
import vfsgpu as V


def uniform_box(N, all_periodic=False):
    Ni, Nj, Nk = N
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=1 if all_periodic else 0, k_periodic=1,
               bctype=[7] * 6 if all_periodic else [7, 7, 1, 1, 7, 7], immersed=0, les=0, second_order=0,
               dpdz_set=1, ren=1000.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0)
    coor = V.channel_grid(Ni + 1, Nj + 1, Nk + 1, W.LX, W.LY, W.LZ, 0.0, 0.0)
    return cfg, coor, Oracle(cfg, coor)


@pytest.mark.parametrize("df,tol", [(0, 1e-15), (10, 1e-14), (8, 1e-8)])
def test_delta_partition_of_unity_and_first_moment(df, tol):
    """sum_m phi(x - m) = 1 and sum_m (x - m) phi(x - m) = 0 for any shift x
    (dfunc_s4h carries the reference's truncated pi, hence its 1e-8)."""
    for x in np.random.default_rng(df).uniform(0, 1, 25):
        m = np.arange(-4, 6)
        phi = np.array([dfunc(df, x - q) for q in m])
        assert abs(phi.sum() - 1) < tol
        assert abs(((x - m) * phi).sum()) < tol


def test_delta_closed_forms():
    assert dfunc(0, 0.25) == 0.75 and dfunc(0, -1.0) == 0.0
    assert dfunc(6, 1.0, 4.0) == 3.0 / 16.0 and dfunc(6, 5.0, 4.0) == 0.0
    assert dfunc(7, 0.0, 2.0) == pytest.approx(1 / (2 * np.sqrt(3.1415926)), rel=1e-15)
    assert dfunc(10, 0.0) == pytest.approx((np.pi + 2 * np.sin(np.pi / 4) * 2) / (4 * np.pi), rel=1e-15)
    assert dfunc(10, 2.5) == pytest.approx(0.0, abs=1e-16) and dfunc(10, 3.0) == 0.0
    assert dfunc(8, 0.0) == pytest.approx(3 / 8 + 3.14159265 / 32, rel=1e-15)


def _element_case(o, coor, rng, n=3):
    """n elements well inside the grid with random forces and areas."""
    c = o.g["cent"].reshape(o.mz, o.my, o.mx, 3)
    pts = []
    for _ in range(n):
        k, j, i = (rng.integers(7, m - 8) for m in (o.mz, o.my, o.mx))
        pts.append(c[k, j, i] + rng.uniform(-0.3, 0.3, 3) * (c[k + 1, j + 1, i + 1] - c[k, j, i]))
    cent = np.array(pts).T.copy()
    F = rng.standard_normal((3, n))
    dA = rng.uniform(0.5, 2.0, n)
    box, dh = o.rotor_elements(coor, cent)
    return cent, F, dA, box, dh


@pytest.mark.parametrize("df", [0, 8, 10])
def test_uniform_grid_conservation_and_centroid(df):
    N = (24, 22, 20)
    cfg, coor, o = uniform_box(N)
    h = np.array([W.LX / N[0], W.LY / N[1], W.LZ / N[2]])
    rng = np.random.default_rng(3 + df)
    cent, F, dA, box, dh = _element_case(o, coor, rng, n=1)
    assert np.allclose(dh[:, 0], h, rtol=1e-12)
    f = o.calc_f_eul(coor, cent, F, dA, box, df=df).reshape(o.mz, o.my, o.mx, 3)
    tol = 1e-12 if df != 8 else 1e-7
    co = coor.reshape(o.mz, o.my, o.mx, 3)
    for q in range(3):
        assert abs(f[..., q].sum() * h[q] - F[q, 0] * dA[0]) < tol * abs(F[q, 0] * dA[0])
    # face centres: the i face of cell (k,j,i) averages nodes (j, j-1) x (k, k-1)
    xi = 0.25 * (co[1:, 1:, :, 0] + co[:-1, 1:, :, 0] + co[1:, :-1, :, 0] + co[:-1, :-1, :, 0])
    yj = 0.25 * (co[1:, :, 1:, 1] + co[:-1, :, 1:, 1] + co[1:, :, :-1, 1] + co[:-1, :, :-1, 1])
    zk = 0.25 * (co[:, 1:, 1:, 2] + co[:, :-1, 1:, 2] + co[:, 1:, :-1, 2] + co[:, :-1, :-1, 2])
    fx, fy, fz = f[1:, 1:, :, 0], f[1:, :, 1:, 1], f[:, 1:, 1:, 2]
    assert abs((fx * xi).sum() / fx.sum() - cent[0, 0]) < 10 * tol
    assert abs((fy * yj).sum() / fy.sum() - cent[1, 0]) < 10 * tol
    assert abs((fz * zk).sum() / fz.sum() - cent[2, 0]) < 10 * tol


def _rotation(seed):
    q = np.random.default_rng(seed).standard_normal(4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def test_rotation_invariance_on_the_curvilinear_grid():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (40, 30, 28))
    rng = np.random.default_rng(21)
    o.nvert[:] = 0
    cent, F, dA, box, dh = _element_case(o, coor, rng, n=4)
    f1 = o.calc_f_eul(coor, cent, F, dA, box)
    R = _rotation(5)
    c2 = (coor.reshape(-1, 3) @ R.T).reshape(-1)
    o2 = Oracle(cfg, c2)
    box2, dh2 = o2.rotor_elements(c2, R @ cent)
    assert np.array_equal(box2, box)
    assert rel_err(dh2, dh) < 1e-12
    f2 = o2.calc_f_eul(c2, R @ cent, R @ F, dA, box2)
    assert np.count_nonzero(f1) > 1000
    assert rel_err(f2, f1) < 1e-10


def test_element_search_is_the_first_nearest_node():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (16, 12, 14))
    rng = np.random.default_rng(8)
    co = coor.reshape(o.mz, o.my, o.mx, 3)
    n = 40
    lo = co[1:-1, 1:-1, 1:-1].reshape(-1, 3).min(0)
    hi = co[1:-1, 1:-1, 1:-1].reshape(-1, 3).max(0)
    cent = rng.uniform(lo, hi, (n, 3)).T.copy()
    cent[:, :4] = co[5, 4, 3][:, None]                   # exactly on a node
    box, dh = o.rotor_elements(coor, cent)
    for l in range(n):
        best, arg = 1e13, None
        for i in range(1, o.mx - 1):
            for j in range(1, o.my - 1):
                for k in range(1, o.mz - 1):
                    d = np.sqrt(((cent[:, l] - co[k, j, i]) ** 2).sum())
                    if d < best:
                        best, arg = d, (i, j, k)
        for q in range(3):
            m = (o.mx, o.my, o.mz)[q]
            assert box[2 * q, l] == max(arg[q] - 6, 1) and box[2 * q + 1, l] == min(arg[q] + 6, m - 1)
        a, b, e = (o.g[nm].reshape(-1, 3)[(arg[2] * o.my + arg[1]) * o.mx + arg[0]] for nm in ("csi", "eta", "zet"))
        aj = o.g["aj"][(arg[2] * o.my + arg[1]) * o.mx + arg[0]]
        assert dh[0, l] == 1.0 / aj / np.sqrt(a @ a) and dh[2, l] == 1.0 / aj / np.sqrt(e @ e)


def test_mask_rule():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (40, 30, 28))
    rng = np.random.default_rng(13)
    nv = o.nvert.copy()
    # elements around the sphere surface
    cent = (np.array(W.SPHERE[:3])[:, None] + 0.27 * rng.standard_normal((3, 30)) / np.sqrt(3)).copy()
    F = rng.standard_normal((3, 30))
    dA = np.ones(30)
    box, dh = o.rotor_elements(coor, cent)
    f_mask = o.calc_f_eul(coor, cent, F, dA, box).reshape(-1, 3)
    o.nvert[:] = 0
    f_free = o.calc_f_eul(coor, cent, F, dA, box).reshape(-1, 3)
    o.nvert[:] = nv
    a = nv.reshape(o.mz, o.my, o.mx)
    s = np.zeros_like(a)
    inner = a[1:-1, 1:-1, 1:-1]
    pad = np.pad(inner, 1)
    for dk in range(3):
        for dj in range(3):
            for di in range(3):
                s[1:-1, 1:-1, 1:-1] += pad[dk:dk + inner.shape[0], dj:dj + inner.shape[1], di:di + inner.shape[2]]
    zero = (s > 2.9).reshape(-1)
    assert zero.any() and np.count_nonzero(f_free[zero]) > 0
    assert not f_mask[zero].any()
    assert np.array_equal(f_mask[~zero], f_free[~zero])


def test_body_force_enters_the_residual_last():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    rng = np.random.default_rng(4)
    f = rng.standard_normal(3 * o.P)
    args = [ucont, ucont * 0.9, rng.standard_normal(3 * o.P), rng.standard_normal(3 * o.P), o.sca() + 1e-3]
    r0, r1 = o.vec(), o.vec()
    o.formfunction_snes_f(*args, None, o.vec(), ucat.copy(), r0)
    o.formfunction_snes_f(*args, f, o.vec(), ucat.copy(), r1)
    assert np.array_equal(r1, r0 + 1. * f)


# ---------------------------------------------------------------------------
# GPU parity (HIP path through the C-ABI against the oracle)
# ---------------------------------------------------------------------------
TOL = 1e-12   # FP64; the delta functions call sin/exp/asin/sqrt


def _ring(n, center, radius, normal_axis=0, seed=0):
    """n elements on a ring (the actuator-disk layout), forces along the axis"""
    rng = np.random.default_rng(seed)
    th = 2 * np.pi * np.arange(n) / n
    e1, e2 = [a for a in range(3) if a != normal_axis]
    cent = np.zeros((3, n))
    cent[normal_axis] = center[normal_axis]
    cent[e1] = center[e1] + radius * np.cos(th)
    cent[e2] = center[e2] + radius * np.sin(th)
    F = np.zeros((3, n))
    F[normal_axis] = -1.0 + 0.1 * rng.standard_normal(n)
    F[e1] = 0.05 * rng.standard_normal(n)
    dA = np.full(n, 2 * np.pi * radius / n * 0.1)
    return cent, F, dA


@pytest.mark.gpu
@pytest.mark.parametrize("name,df,width_mode", [("C2", 10, 0), ("C2", 7, 0), ("C2", 8, 1), ("C1w", 0, 3),
                                                ("C1w", 6, 2)])
def test_gpu_actuator_parity(name, df, width_mode):
    print(f"[actuator] {name} df={df} width_mode={width_mode}", flush=True)
    cfg, spec, coor, o, ucat, ucont = oracle_case(name, (40, 30, 28))
    ctx = gpu_case(cfg, spec, coor)
    assert np.array_equal(ctx.nvert().host(), o.nvert)
    center = (W.LX / 2, 1.0, W.LZ / 2) if name == "C2" else (W.LX / 2, 0.6, W.LZ / 2)
    cent, F, dA = _ring(48, center, 0.4, seed=df)
    if width_mode in (2, 3):   # widths from the element area: about two cells
        dA = np.full(dA.shape, 0.3 if width_mode == 2 else 0.09)
    box, dh = o.rotor_elements(coor, cent)
    dcoor = V.DeviceArray.from_host(coor)
    gbox, gdh = ctx.rotor_elements(dcoor, V.DeviceArray.from_host(cent.reshape(-1)))
    assert np.array_equal(gbox.host().reshape(6, -1), box)
    assert np.array_equal(gdh.host().reshape(3, -1), dh)
    fixed = (0.2, 0.1, 0.15)
    want = o.calc_f_eul(coor, cent, F, dA, box, df=df, halfwidth=2.0, width_mode=width_mode, dh_fixed=fixed)
    f = ctx.vec()
    ctx.calc_f_eul(dcoor, V.DeviceArray.from_host(cent.reshape(-1)), V.DeviceArray.from_host(F.reshape(-1)),
                   V.DeviceArray.from_host(dA), gbox, f, deltafunc=df, halfwidth=2.0, width_mode=width_mode,
                   dh_fixed=fixed)
    got = f.host()
    assert np.count_nonzero(want) > 100
    assert rel_err(got, want) < TOL
    assert np.array_equal(got == 0, want == 0)
    ctx.close()


@pytest.mark.gpu
def test_gpu_body_force_in_the_residual():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    ctx = gpu_case(cfg, spec, coor)
    rng = np.random.default_rng(9)
    f = rng.standard_normal(3 * o.P)
    U_o, rhs_o, dP = 0.9 * ucont, rng.standard_normal(3 * o.P), rng.standard_normal(3 * o.P)
    nut = o.sca() + 1e-3
    want = o.vec()
    o.formfunction_snes_f(ucont, U_o, rhs_o, dP, nut, f, o.vec(), ucat.copy(), want)
    dv = {k: V.DeviceArray.from_host(v) for k, v in dict(U=ucont, U_o=U_o, rhs_o=rhs_o, dP=dP, nut=nut, f=f,
                                                        ucat=ucat).items()}
    ctx.set_body_force(dv["f"])
    out, lu = ctx.vec(), ctx.vec()
    ctx.formfunction_snes(dv["U"], dv["U_o"], dv["rhs_o"], dv["dP"], dv["nut"], lu, dv["ucat"], out)
    assert rel_err(out.host(), want) < TOL
    ctx.set_body_force(None)
    ctx.formfunction_snes(dv["U"], dv["U_o"], dv["rhs_o"], dv["dP"], dv["nut"], lu, V.DeviceArray.from_host(ucat), out)
    plain = o.vec()
    o.formfunction_snes_f(ucont, U_o, rhs_o, dP, nut, None, o.vec(), ucat.copy(), plain)
    assert rel_err(out.host(), plain) < TOL
    ctx.close()


@pytest.mark.gpu
def test_gpu_element_boxes_outside_the_interior_are_rejected():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    ctx = gpu_case(cfg, spec, coor)
    box = np.array([[0], [3], [1], [3], [1], [3]], dtype=np.int32)
    one = V.DeviceArray.from_host(np.zeros(3))
    with pytest.raises(V.VfsError, match="interior"):
        ctx.calc_f_eul(V.DeviceArray.from_host(coor), one, one, V.DeviceArray.from_host(np.ones(1)),
                       V.DeviceArray.from_host(box.reshape(-1)), ctx.vec())
    ctx.close()
