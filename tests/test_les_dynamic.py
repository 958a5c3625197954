"""The dynamic Smagorinsky procedure (Compute_Smagorinsky_Constant_1 with
-les 2, les.c:67-1133) and the eddy viscosity with a per-cell Cs.

CPU tests pin the oracle:
  - on a uniform Cartesian triply periodic box the reference's contravariant
    contractions reduce to the textbook Germano-Lilly model (Cartesian
    L_ij M_ij, M_ij M_ij, test filter width 2 Delta), written here with numpy
    rolls and separable (1, 2, 1)/4 box filters;
  - the model is a scalar: rotating the grid and the velocity rigidly leaves
    Cs unchanged on the curvilinear channel with an immersed sphere (a
    transposed index in the contractions breaks this);
  - integrate_testfilter against its closed form and hand-computed weights;
  - the Cs-averaging filters on constant fields, with the reference's j rule
    (the j = my-1 ghost plane keeps its weight), the clipping and the
    fresh-start and -les 1 rules.
GPU tests compare the HIP path with the oracle on the parity cases.
"""
import numpy as np
import pytest

from cases import SMALL, gpu_case, oracle_case
from conftest import rel_err
from oracle_py import Oracle, testfilter as integrate_testfilter
from vfsgpu import workloads as W
import vfsgpu as V

LES_EPS = 1e-4


def uniform_box(N, all_periodic=True, clark=0):
    Ni, Nj, Nk = N
    cfg = dict(mx=Ni + 2, my=Nj + 2, mz=Nk + 2, i_periodic=1, j_periodic=1 if all_periodic else 0, k_periodic=1,
               bctype=[7] * 6 if all_periodic else [7, 7, 1, 1, 7, 7], immersed=0, les=1, second_order=0,
               dpdz_set=1, ren=1000.0, dt=1e-3, st=1.0, mean_pressure_gradient=0.0, clark=clark)
    coor = V.channel_grid(Ni + 1, Nj + 1, Nk + 1, W.LX, W.LY, W.LZ, 0.0, 0.0)
    return cfg, Oracle(cfg, coor)


def periodic_ghosts(o, v, comps):
    a = v.reshape(o.mz, o.my, o.mx, comps)
    a[:, :, 0] = a[:, :, -2]; a[:, :, -1] = a[:, :, 1]
    a[:, 0] = a[:, -2]; a[:, -1] = a[:, 1]
    a[0] = a[-2]; a[-1] = a[1]
    return v


def box(f):
    """the uniform-weight integrate_testfilter: (1, 2, 1)/4 along each axis"""
    for ax in (-1, -2, -3):
        f = 0.25 * np.roll(f, 1, ax) + 0.5 * f + 0.25 * np.roll(f, -1, ax)
    return f


def textbook_germano_lilly(u, h, max_cs, clark=False):
    """u: (3, Nk, Nj, Ni) periodic; h = (dx, dy, dz).  Returns LM, MM, Cs.
    clark: the mixed model's Germano term, L - N with N = 4 R(u^) - <T(u)>
    and the Clark stress T_ab = sum_m d_m u_a d_m u_b h_m^2 / 12."""
    axes = (-1, -2, -3)

    def grad(v):
        return np.array([[(np.roll(v[a], -1, axes[b]) - np.roll(v[a], 1, axes[b])) / (2 * h[b]) for b in range(3)]
                         for a in range(3)])

    def strain(v):
        D = grad(v)
        S = 0.5 * (D + D.transpose(1, 0, 2, 3, 4))
        return S, np.sqrt(2 * (S * S).sum(axis=(0, 1)))

    vol = h[0] * h[1] * h[2]
    S, Sa = strain(u)
    uf = np.array([box(u[a]) for a in range(3)])
    L = np.array([[box(u[a] * u[b]) - uf[a] * uf[b] for b in range(3)] for a in range(3)])
    Sh, Sha = strain(uf)
    M = np.array([[-(8 * vol) ** (2 / 3) * Sha * Sh[a, b] + vol ** (2 / 3) * box(Sa * S[a, b]) for b in range(3)]
                  for a in range(3)])
    if clark:
        def clark_stress(v):
            D = grad(v)
            return np.array([[sum(D[a, m] * D[b, m] * h[m] ** 2 for m in range(3)) / 12 for b in range(3)]
                             for a in range(3)])
        T = clark_stress(u)
        N = 4 * clark_stress(uf) - np.array([[box(T[a, b]) for b in range(3)] for a in range(3)])
        L = L - N
    LM = (L * M).sum(axis=(0, 1))
    MM = (M * M).sum(axis=(0, 1))
    C = 0.5 * box(LM) / (box(MM) + LES_EPS)
    C = np.minimum(np.maximum(C, 0), max_cs)
    return LM, MM, box(C)


def test_testfilter_closed_form_and_weights():
    rng = np.random.default_rng(3)
    v = rng.standard_normal((3, 3, 3))
    # uniform weights: the separable (1, 2, 1)/4 kernel
    k1 = np.array([0.25, 0.5, 0.25])
    kern = k1[:, None, None] * k1[None, :, None] * k1[None, None, :]
    assert abs(integrate_testfilter(v, np.ones((3, 3, 3))) - (kern * v).sum()) < 1e-15
    # general weights: each point counts once per octant that holds it
    w = rng.uniform(0.1, 2.0, (3, 3, 3))
    cnt = np.array([1, 2, 1])
    mult = cnt[:, None, None] * cnt[None, :, None] * cnt[None, None, :]
    want = (mult * v * w).sum() / (mult * w).sum()
    assert abs(integrate_testfilter(v, w) - want) < 1e-14 * abs(want) + 1e-15
    # a zero weight removes that point
    w[2, 1, 0] = 0
    want = (mult * v * w).sum() / (mult * w).sum()
    assert abs(integrate_testfilter(v, w) - want) < 1e-14 * abs(want) + 1e-15


@pytest.mark.parametrize("clark", [0, 1], ids=["smagorinsky", "clark"])
@pytest.mark.parametrize("N", [(10, 8, 12), (9, 11, 7)])
def test_reduces_to_textbook_germano_lilly_on_a_uniform_box(N, clark):
    """Triply periodic uniform box with unequal spacings: the oracle's LM, MM
    (pass 2) and final Cs equal the Cartesian textbook model, with and
    without the Clark term (-clark, les.c:648)."""
    cfg, o = uniform_box(N, clark=clark)
    Ni, Nj, Nk = N
    h = (W.LX / Ni, W.LY / Nj, W.LZ / Nk)
    rng = np.random.default_rng(7)
    # a smooth field (low Fourier modes) plus a little noise: Cs takes values
    # inside (0, max_cs) as well as on both clips
    cent = o.g["cent"].reshape(o.mz, o.my, o.mx, 3)[1:-1, 1:-1, 1:-1]
    x, y, z = cent[..., 0], cent[..., 1], cent[..., 2]
    u = np.zeros((3, Nk, Nj, Ni))
    for a in range(3):
        for _ in range(4):
            kx, ky, kz = rng.integers(0, 3, 3)
            ph = rng.uniform(0, 2 * np.pi)
            u[a] += rng.standard_normal() * np.cos(kx * x + ky * np.pi * y + 2 * kz * z + ph)
        u[a] += 0.2 * rng.standard_normal(x.shape)
    ucat = np.zeros(3 * o.P)
    ucat.reshape(o.mz, o.my, o.mx, 3)[1:-1, 1:-1, 1:-1] = np.moveaxis(u, 0, -1)
    periodic_ghosts(o, ucat, 3)
    LM_t, MM_t, Cs_t = textbook_germano_lilly(u, h, 0.5, clark=bool(clark))
    if clark:   # the Clark term moves LM by a share of its size
        LM_s = textbook_germano_lilly(u, h, 0.5)[0]
        assert np.abs(LM_t - LM_s).max() > 0.05 * np.abs(LM_s).max()
    grad, sabs, uf = o.dsm_pass1(ucat)
    LM, MM = o.dsm_pass2(ucat, grad, sabs, uf)
    inner = (slice(1, -1),) * 3
    assert rel_err(LM.reshape(o.mz, o.my, o.mx)[inner], LM_t) < 1e-11
    assert rel_err(MM.reshape(o.mz, o.my, o.mx)[inner], MM_t) < 1e-11
    cs = o.smagorinsky_constant(ucat)
    got = cs.reshape(o.mz, o.my, o.mx)[inner]
    assert rel_err(got, Cs_t) < 1e-11
    # the case exercises the model, not only its clips
    raw = 0.5 * box(LM_t) / (box(MM_t) + LES_EPS)
    assert (raw < 0).mean() > 0.05 and ((raw > 0) & (raw < 0.5)).mean() > 0.3
    # nu_t = Cs |S| Delta^2
    nut = o.sca()
    o.eddy_viscosity_cs(ucat, cs, nut)
    vol = h[0] * h[1] * h[2]
    assert rel_err(nut.reshape(o.mz, o.my, o.mx)[inner], got * vol ** (2 / 3) * sabs.reshape(o.mz, o.my, o.mx)[inner]) < 1e-13


def _rotation(seed):
    q = np.random.default_rng(seed).standard_normal(4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


@pytest.mark.parametrize("name", ["C2", "C1w"])
def test_rotation_invariance_on_curvilinear_grids(name):
    """Rigidly rotated grid and velocity (same mask): same Cs and nu_t."""
    cfg, spec, coor, o, ucat, ucont = oracle_case(name)
    R = _rotation(11)
    assert abs(np.linalg.det(R) - 1) < 1e-14
    o2 = Oracle(cfg, (coor.reshape(-1, 3) @ R.T).reshape(-1))
    o2.nvert[:] = o.nvert
    u2 = (ucat.reshape(-1, 3) @ R.T).reshape(-1)
    cs1, cs2 = o.smagorinsky_constant(ucat), o2.smagorinsky_constant(u2)
    assert np.max(cs1) > 0.01 and np.count_nonzero(cs1) > cs1.size // 4
    assert rel_err(cs2, cs1) < 1e-9
    n1, n2 = o.sca(), o2.sca()
    o.eddy_viscosity_cs(ucat, cs1, n1)
    o2.eddy_viscosity_cs(u2, cs2, n2)
    assert rel_err(n2, n1) < 1e-9
    # pass 1: the filtered velocity rotates with the field, |S| does not change
    g1, s1, f1 = o.dsm_pass1(ucat)
    g2, s2, f2 = o2.dsm_pass1(u2)
    assert rel_err((f2.reshape(-1, 3) @ R), f1.reshape(-1, 3)) < 1e-12
    assert rel_err(s2, s1) < 1e-10


def test_cs_filters_and_the_j_rule_on_constant_fields():
    """Uniform channel (j walls), no solids.  The smoothing of a constant c:
    c everywhere except the top row j = my-2, whose box still counts the
    j = my-1 ghost plane (value 0, a quarter of the weight): 3c/4.  Pass 3
    of constant LM, MM: 0.5 LM / (MM + eps), with the top row seeing 3/4 of
    both."""
    cfg, o = uniform_box((8, 9, 7), all_periodic=False)
    c = 0.123
    cs = o.sca()
    a = cs.reshape(o.mz, o.my, o.mx)
    a[1:-1, 1:-1, 1:-1] = c
    periodic_ghosts(o, cs, 1)
    a[:, 0] = 0; a[:, -1] = 0
    o.dsm_smooth(cs)
    inner = a[1:-1, 1:-1, 1:-1]
    assert np.allclose(inner[:, :-1], c, rtol=1e-14, atol=0)
    assert np.allclose(inner[:, -1], 0.75 * c, rtol=1e-14, atol=0)
    lm, mm = 0.02, 0.3
    LM, MM = o.sca() + lm, o.sca() + mm
    LMg, MMg = LM.reshape(o.mz, o.my, o.mx), MM.reshape(o.mz, o.my, o.mx)
    LMg[:, -1] = 0; MMg[:, -1] = 0   # the ghosts pass 2 leaves at zero
    out = o.dsm_pass3(LM, MM).reshape(o.mz, o.my, o.mx)
    assert np.allclose(out[1:-1, 1:-2, 1:-1], 0.5 * lm / (mm + LES_EPS), rtol=1e-14, atol=0)
    assert np.allclose(out[1:-1, -2, 1:-1], 0.5 * 0.75 * lm / (0.75 * mm + LES_EPS), rtol=1e-14, atol=0)
    # ghosts: zero on the walls, periodic copies on the i and k planes
    assert not out[:, 0].any() and not out[:, -1].any()
    assert np.array_equal(out[:, :, 0], out[:, :, -2]) and np.array_equal(out[0], out[-2])


def test_clipping_rules():
    """Negative C -> 0, above max_cs -> max_cs, IB nodes at least wall_cs,
    solids 0."""
    cfg, o = uniform_box((8, 9, 7), all_periodic=False)
    nv = o.nvert.reshape(o.mz, o.my, o.mx)
    nv[3, 4, 4] = 3.0           # a solid cell
    nv[3, 4, 5] = 1.0           # an IB node next to it
    for lm, want_f, want_ib in [(-0.01, 0.0, 0.001), (1e-5, None, 0.001), (5.0, 0.2, 0.2)]:
        LM, MM = o.sca() + lm, o.sca() + 0.3
        out = o.dsm_pass3(LM, MM, max_cs=0.2).reshape(o.mz, o.my, o.mx)
        assert out[3, 4, 4] == 0
        assert out[3, 4, 5] == want_ib
        f = out[4, 3, 2]
        assert f == pytest.approx(want_f if want_f is not None else 0.5 * 1e-5 / (0.3 + LES_EPS), rel=1e-14, abs=0)


def test_fresh_start_and_constant_rules():
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2")
    assert not o.smagorinsky_constant(ucat, ti=1).any()
    assert not o.smagorinsky_constant(ucat, ti=0).any()
    assert o.smagorinsky_constant(ucat, ti=1, rstart=1).any()    # a restart computes it
    assert o.smagorinsky_constant(ucat, ti=1, tistart=5).any()
    assert np.all(o.smagorinsky_constant(ucat, les=1) == 0.01)
    cs = o.smagorinsky_constant(ucat)
    nv = o.nvert
    assert not cs[nv > 1.1].any()                              # solids
    assert np.all(cs[(nv > 0.1) & (nv < 1.1)] >= 0.001)        # IB nodes
    assert cs.min() >= 0 and cs.max() <= 0.5
    # nu_t with Cs = 0.01 everywhere equals the constant-Cs path
    n1, n2 = o.sca(), o.sca()
    o.eddy_viscosity(ucat, 0.01, n1)
    o.eddy_viscosity_cs(ucat, np.full(o.P, 0.01), n2)
    assert np.array_equal(n1, n2)


# ---------------------------------------------------------------------------
# GPU parity (HIP path through the C-ABI against the oracle)
# ---------------------------------------------------------------------------
TOL = 1e-12   # FP64; the two paths differ only through pow(x, 1/3)


def _ucat_after_contra2cart(name):
    cfg, spec, coor, o, ucat, ucont = oracle_case(name)
    uc, u = ucont.copy(), ucat.copy()
    o.contra2cart(uc, u)
    return cfg, spec, coor, o, u


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["C1", "C2", "C1w", "C2w"])
def test_gpu_dynamic_smagorinsky_parity(name):
    print(f"[dsm] {name} {SMALL[name]}", flush=True)
    cfg, spec, coor, o, u = _ucat_after_contra2cart(name)
    ctx = gpu_case(cfg, spec, coor)
    assert np.array_equal(ctx.nvert().host(), o.nvert)
    du = V.DeviceArray.from_host(u)
    want = o.smagorinsky_constant(u, max_cs=0.3)
    cs = ctx.sca()
    ctx.smagorinsky_constant(du, cs, max_cs=0.3)
    got = cs.host()
    assert rel_err(got, want) < TOL
    assert np.array_equal(got == 0, want == 0)          # the zeros (solids, ghosts, clips) match
    assert np.array_equal(got == 0.3, want == 0.3)
    assert np.count_nonzero(want) > want.size // 4
    # a second call on the same context reuses its scratch
    ctx.smagorinsky_constant(du, cs, max_cs=0.3)
    assert np.array_equal(cs.host(), got)
    nut_o, nut = o.sca(), ctx.sca()
    o.eddy_viscosity_cs(u, want, nut_o)
    ctx.eddy_viscosity_cs(du, cs, nut)
    assert rel_err(nut.host(), nut_o) < TOL
    ctx.close()


@pytest.mark.gpu
def test_gpu_dynamic_smagorinsky_periodic_box():
    """The triply periodic uniform box of the textbook check, on the GPU."""
    N = (10, 8, 12)
    cfg, o = uniform_box(N)
    rng = np.random.default_rng(5)
    ucat = np.zeros(3 * o.P)
    ucat.reshape(o.mz, o.my, o.mx, 3)[1:-1, 1:-1, 1:-1] = rng.standard_normal((N[2], N[1], N[0], 3))
    periodic_ghosts(o, ucat, 3)
    ctx = V.Ctx(cfg)
    Ni, Nj, Nk = N
    ctx.build_metrics(V.DeviceArray.from_host(V.channel_grid(Ni + 1, Nj + 1, Nk + 1, W.LX, W.LY, W.LZ, 0.0, 0.0)))
    cs = ctx.sca()
    ctx.smagorinsky_constant(V.DeviceArray.from_host(ucat), cs)
    assert rel_err(cs.host(), o.smagorinsky_constant(ucat)) < TOL
    ctx.close()


@pytest.mark.gpu
def test_gpu_fresh_start_and_constant_rules():
    cfg, spec, coor, o, u = _ucat_after_contra2cart("C2")
    ctx = gpu_case(cfg, spec, coor)
    du, cs = V.DeviceArray.from_host(u), ctx.sca()
    ctx.smagorinsky_constant(du, cs, les=1)
    assert np.all(cs.host() == 0.01)
    ctx.smagorinsky_constant(du, cs, ti=1)
    assert not cs.host().any()
    ctx.smagorinsky_constant(du, cs, les=1, ti=0)
    assert not cs.host().any()
    ctx.smagorinsky_constant(du, cs, ti=1, rstart=1)
    assert rel_err(cs.host(), o.smagorinsky_constant(u, ti=1, rstart=1)) < TOL
    with pytest.raises(V.VfsError, match="les must be"):
        ctx.smagorinsky_constant(du, cs, les=4)
    with pytest.raises(V.VfsError, match="homogeneous"):
        ctx.smagorinsky_constant(du, cs, i_homo=1, j_homo=1)
    ctx.close()


OPTION_CASES = [dict(ik=1), dict(i_homo=1, k_homo=1), dict(i_homo=1), dict(j_homo=1), dict(k_homo=1),
                dict(les=3, ti=50), dict(les=3, ti=150), dict(les=3, ti=5, rstart=1, ik=1)]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["C2", "C1w"])
@pytest.mark.parametrize("opt", OPTION_CASES, ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()))
def test_gpu_dynamic_options(name, opt):
    print(f"[dsm options] {name} {opt}", flush=True)
    cfg, spec, coor, o, u = _ucat_after_contra2cart(name)
    ctx = gpu_case(cfg, spec, coor)
    du = V.DeviceArray.from_host(u)
    prev = np.abs(np.random.default_rng(3).standard_normal(o.P)) * 0.05
    want = o.smagorinsky_constant(u, cs_prev=prev, max_cs=0.3, **opt)
    cs = V.DeviceArray.from_host(prev)
    ctx.smagorinsky_constant(du, cs, max_cs=0.3, **opt)
    got = cs.host()
    assert np.count_nonzero(want) > want.size // 4
    assert rel_err(got, want) < TOL
    assert np.array_equal(got == 0, want == 0)
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["C2", "C1w"])
@pytest.mark.parametrize("opt", [dict(), dict(ik=1), dict(les=3, ti=150), dict(i_homo=1, k_homo=1)],
                         ids=["box", "ik", "les3", "ik-homo"])
def test_gpu_dynamic_clark(name, opt):
    """-clark: the Clark term of the Germano identity (les.c:648) on the
    parity cases, with the box and the i-k plane filters."""
    print(f"[dsm clark] {name} {opt}", flush=True)
    N = SMALL[name]
    cfg, spec = W.make_config(name, N)
    cfg["clark"] = 1
    coor = W.grid(spec)
    o = Oracle(cfg, coor)
    if spec["immersed"]:
        o.nvert_sphere(*W.SPHERE)
    o.wall_mask()
    metrics = {m: o.g[m] for m in ("cent", "icsi", "jeta", "kzet")}
    ucat, ucont = W.fields(cfg, spec, metrics, o.nvert)
    o.contra2cart(ucont, ucat)
    ctx = gpu_case(cfg, spec, coor)
    assert np.array_equal(ctx.nvert().host(), o.nvert)
    du = V.DeviceArray.from_host(ucat)
    prev = np.abs(np.random.default_rng(4).standard_normal(o.P)) * 0.05
    want = o.smagorinsky_constant(ucat, cs_prev=prev, max_cs=0.3, **opt)
    # the term changes the model's answer on this field (by 10-70 %; under
    # -les 3 the relaxation towards the previous Cs keeps 1e-5 to 1e-3 of it)
    o.c.clark = 0
    assert rel_err(o.smagorinsky_constant(ucat, cs_prev=prev, max_cs=0.3, **opt), want) > 1e-5
    cs = V.DeviceArray.from_host(prev)
    ctx.smagorinsky_constant(du, cs, max_cs=0.3, **opt)
    got = cs.host()
    assert np.count_nonzero(want) > want.size // 4
    assert rel_err(got, want) < TOL
    assert np.array_equal(got == 0, want == 0)
    ctx.close()


# ---------------------------------------------------------------------------
# the other options of the routine: -testfilter_ik, -i/j/k_homo_filter, -les 3
# ---------------------------------------------------------------------------
from oracle_py import testfilter_ik as simpson_ik  # noqa: E402


def test_testfilter_ik_is_simpson_in_the_i_k_plane():
    """x^2 (i offset squared) filters to 1/3, k^2 likewise, x k to 0; the
    j = -1 and +1 planes and the weights are ignored."""
    R, Q, P = np.meshgrid(np.arange(3) - 1.0, np.arange(3) - 1.0, np.arange(3) - 1.0, indexing="ij")
    assert simpson_ik(P ** 2) == pytest.approx(1 / 3, rel=1e-15)
    assert simpson_ik(R ** 2) == pytest.approx(1 / 3, rel=1e-15)
    assert simpson_ik(R * P) == 0.0
    v = np.random.default_rng(2).standard_normal((3, 3, 3))
    v2 = v.copy()
    v2[:, 0] += 5.0
    v2[:, 2] -= 3.0
    assert simpson_ik(v2) == simpson_ik(v)


def simpson_xz(f):
    for ax in (-1, -3):
        f = (np.roll(f, 1, ax) + 4 * f + np.roll(f, -1, ax)) / 6
    return f


@pytest.mark.parametrize("clark", [0, 1], ids=["smagorinsky", "clark"])
def test_testfilter_ik_reduces_to_the_textbook_2d_filter(clark):
    """Triply periodic uniform box: with -testfilter_ik the model is Germano-
    Lilly with the 2-D Simpson filter in x-z and test width 5^(1/3) Delta,
    and the Clark term's <T> takes the same filter."""
    N = (10, 8, 12)
    cfg, o = uniform_box(N, clark=clark)
    h = (W.LX / N[0], W.LY / N[1], W.LZ / N[2])
    rng = np.random.default_rng(17)
    u = rng.standard_normal((3, N[2], N[1], N[0]))
    ucat = np.zeros(3 * o.P)
    ucat.reshape(o.mz, o.my, o.mx, 3)[1:-1, 1:-1, 1:-1] = np.moveaxis(u, 0, -1)
    periodic_ghosts(o, ucat, 3)
    axes = (-1, -2, -3)

    def strain(v):
        D = np.array([[(np.roll(v[a], -1, axes[b]) - np.roll(v[a], 1, axes[b])) / (2 * h[b]) for b in range(3)]
                      for a in range(3)])
        S = 0.5 * (D + D.transpose(1, 0, 2, 3, 4))
        return S, np.sqrt(2 * (S * S).sum(axis=(0, 1)))
    vol = h[0] * h[1] * h[2]
    S, Sa = strain(u)
    uf = np.array([simpson_xz(u[a]) for a in range(3)])
    L = np.array([[simpson_xz(u[a] * u[b]) - uf[a] * uf[b] for b in range(3)] for a in range(3)])
    Sh, Sha = strain(uf)
    M = np.array([[-5 ** (2 / 3) * vol ** (2 / 3) * Sha * Sh[a, b] + vol ** (2 / 3) * simpson_xz(Sa * S[a, b])
                   for b in range(3)] for a in range(3)])
    if clark:
        def clark_stress(v):
            D = np.array([[(np.roll(v[a], -1, axes[m]) - np.roll(v[a], 1, axes[m])) / (2 * h[m]) for m in range(3)]
                          for a in range(3)])
            return np.array([[sum(D[a, m] * D[b, m] * h[m] ** 2 for m in range(3)) / 12 for b in range(3)]
                             for a in range(3)])
        T = clark_stress(u)
        L = L - (4 * clark_stress(uf) - np.array([[simpson_xz(T[a, b]) for b in range(3)] for a in range(3)]))
    LM_t, MM_t = (L * M).sum(axis=(0, 1)), (M * M).sum(axis=(0, 1))
    grad, sabs, ufo = o.dsm_pass1(ucat, ik=1)
    LM, MM = o.dsm_pass2(ucat, grad, sabs, ufo, ik=1)
    inner = (slice(1, -1),) * 3
    assert rel_err(LM.reshape(o.mz, o.my, o.mx)[inner], LM_t) < 1e-11
    assert rel_err(MM.reshape(o.mz, o.my, o.mx)[inner], MM_t) < 1e-11
    C = np.minimum(np.maximum(0.5 * simpson_xz(LM_t) / (simpson_xz(MM_t) + LES_EPS), 0), 0.5)
    cs = o.smagorinsky_constant(ucat, ik=1)
    assert rel_err(cs.reshape(o.mz, o.my, o.mx)[inner], simpson_xz(C)) < 1e-11


@pytest.mark.parametrize("mode", ["ik", "i", "j", "k"])
def test_homogeneous_averages(mode):
    """Plane (i and k) or line averages of LM and MM over the fluid cells,
    Cs = 0.5 <LM> / (<MM> + eps) then clipped; no smoothing afterwards."""
    cfg, spec, coor, o, ucat, ucont = oracle_case("C2", (24, 18, 16))
    uc, u = ucont.copy(), ucat.copy()
    o.contra2cart(uc, u)
    flags = dict(i_homo=int("i" in mode), j_homo=int(mode == "j"), k_homo=int("k" in mode))
    grad, sabs, uf = o.dsm_pass1(u, **flags)
    LM, MM = o.dsm_pass2(u, grad, sabs, uf, **flags)
    cs = o.smagorinsky_constant(u, **flags)
    nv = o.nvert.reshape(o.mz, o.my, o.mx)
    L3, M3 = LM.reshape(nv.shape), MM.reshape(nv.shape)
    fl = (nv < 0.1)
    fl[0] = fl[-1] = False; fl[:, 0] = fl[:, -1] = False; fl[:, :, 0] = fl[:, :, -1] = False
    axes = {"ik": (0, 2), "i": (2,), "j": (1,), "k": (0,)}[mode]
    cnt = fl.sum(axis=axes, keepdims=True)
    mL = np.where(cnt > 0, (L3 * fl).sum(axis=axes, keepdims=True) / np.maximum(cnt, 1), 0)
    mM = np.where(cnt > 0, (M3 * fl).sum(axis=axes, keepdims=True) / np.maximum(cnt, 1), 0)
    want = np.broadcast_to(0.5 * mL / (mM + LES_EPS), nv.shape)
    got = cs.reshape(nv.shape)
    inner = np.zeros(nv.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    target = fl if mode != "ik" else inner
    clipped = np.minimum(np.maximum(want, 0), 0.5)
    ibn = (nv > 0.1) & (nv < 1.1) & inner
    clipped = np.where(ibn, np.minimum(np.maximum(np.maximum(0.001, want), 0), 0.5), clipped)
    sel = target & (nv < 1.1)
    assert rel_err(got[sel], clipped[sel]) < 1e-12
    assert not got[nv > 1.1].any()
    if mode != "ik":   # IB nodes keep the pointwise value
        C = np.maximum(0.5 * L3 / (M3 + LES_EPS), 0)
        assert rel_err(got[ibn], np.minimum(np.maximum(np.maximum(0.001, C), 0), 0.5)[ibn]) < 1e-12


def test_les3_relaxation():
    """-les 3: no test filtering of LM, MM and no smoothing; from ti = 100 on
    (or on a restart) Cs = 0.999 Cs_prev + 0.001 C."""
    cfg, o = uniform_box((8, 9, 7), all_periodic=False)
    lm, mm, prev = 0.02, 0.3, 0.2
    LM, MM = o.sca() + lm, o.sca() + mm
    C = 0.5 * lm / (mm + LES_EPS)
    inner = (slice(1, -1),) * 3
    fresh = o.dsm_pass3(LM, MM, cs_prev=o.sca() + prev, les=3, ti=50).reshape(o.mz, o.my, o.mx)
    assert np.all(fresh[inner] == C)        # no filter: the top row is not 3/4 of it
    relax = o.dsm_pass3(LM, MM, cs_prev=o.sca() + prev, les=3, ti=100).reshape(o.mz, o.my, o.mx)
    assert np.all(relax[inner] == (1.0 - 0.001) * prev + 0.001 * C)
    restart = o.dsm_pass3(LM, MM, cs_prev=o.sca() + prev, les=3, ti=5, rstart=1).reshape(o.mz, o.my, o.mx)
    assert np.array_equal(restart, relax)
