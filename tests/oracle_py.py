"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker that GPU results are
compared against; nothing in the product imports this module.
"""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")


class VoCfg(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
                ("i_periodic", C.c_int), ("j_periodic", C.c_int), ("k_periodic", C.c_int),
                ("bctype", C.c_int * 6), ("immersed", C.c_int), ("les", C.c_int),
                ("second_order", C.c_int), ("dpdz_set", C.c_int), ("weno", C.c_int), ("inviscid", C.c_int),
                ("skew", C.c_int), ("clark", C.c_int), ("ren", C.c_double), ("dt", C.c_double), ("st", C.c_double),
                ("mean_pressure_gradient", C.c_double)]


GEOM_FIELDS = ["csi", "eta", "zet", "aj", "icsi", "ieta", "izet", "iaj",
               "jcsi", "jeta", "jzet", "jaj", "kcsi", "keta", "kzet", "kaj", "cent"]
VECTOR_GEOM = {"csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta", "jzet",
               "kcsi", "keta", "kzet", "cent"}


class VoLesOpt(C.Structure):
    _fields_ = [("les", C.c_int), ("i_homo_filter", C.c_int), ("j_homo_filter", C.c_int), ("k_homo_filter", C.c_int),
                ("testfilter_ik", C.c_int), ("max_cs", C.c_double), ("ti", C.c_int), ("tistart", C.c_int),
                ("rstart", C.c_int)]


def les_opt(les=2, ti=5, tistart=0, rstart=0, max_cs=0.5, i_homo=0, j_homo=0, k_homo=0, ik=0):
    return VoLesOpt(les, i_homo, j_homo, k_homo, ik, max_cs, ti, tistart, rstart)


class VoIbOpt(C.Structure):
    _fields_ = [("moving", C.c_int), ("wallfunction", C.c_int), ("ks", C.c_double), ("roughness", C.c_double)]


def _wall_law(wallfunction, roughness):
    return 0 if not wallfunction else (1 if roughness is None else 2)


class VoInflowOpt(C.Structure):
    _fields_ = [("inletprofile", C.c_int), ("reserved", C.c_int), ("inlet_flux", C.c_double),
                ("inlet_area", C.c_double)]


class VoIbSurface(C.Structure):
    _fields_ = [("n_v", C.c_int), ("n_elmt", C.c_int),
                ("x_bp", C.POINTER(C.c_double)), ("y_bp", C.POINTER(C.c_double)), ("z_bp", C.POINTER(C.c_double)),
                ("nv1", C.POINTER(C.c_int)), ("nv2", C.POINTER(C.c_int)), ("nv3", C.POINTER(C.c_int)),
                ("nf_x", C.POINTER(C.c_double)), ("nf_y", C.POINTER(C.c_double)), ("nf_z", C.POINTER(C.c_double)),
                ("u", C.POINTER(C.c_double))]


class IbBody:
    """Host arrays of a triangulated body plus the oracle's view of them."""

    def __init__(self, verts, elems, normals, u=None):
        self.cols = [np.ascontiguousarray(np.asarray(verts, np.float64)[:, q]) for q in range(3)]
        self.cols += [np.ascontiguousarray(np.asarray(elems, np.int32)[:, q]) for q in range(3)]
        self.cols += [np.ascontiguousarray(np.asarray(normals, np.float64)[:, q]) for q in range(3)]
        self.verts, self.elems, self.normals = np.asarray(verts, float), np.asarray(elems), np.asarray(normals, float)
        self.u = None if u is None else np.ascontiguousarray(np.asarray(u, np.float64).ravel())
        self.n_v, self.n_elmt = len(verts), len(elems)
        self.c = VoIbSurface(self.n_v, self.n_elmt, *[ptr(a) for a in self.cols],
                             None if self.u is None else ptr(self.u))


class VoGeom(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in GEOM_FIELDS]


def _build():
    src = os.path.join(ORACLE_DIR, "vfs_oracle.c")
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


_lib = None


def lib():
    global _lib
    if _lib is None:
        _build()
        _lib = _configure(C.CDLL(LIB))
    return _lib


@contextlib.contextmanager
def using(path):
    """Run the enclosed block against another build of the oracle (a
    mutant compiled by the pin tests); every wrapper calls lib()."""
    global _lib
    saved = lib()
    _lib = _configure(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def _configure(L):
    d = C.c_double
    for name in ["vo_min_dt", "vo_divergence", "vo_wf_integrate_F", "vo_wf_find_utau", "vo_testfilter", "vo_testfilter_ik", "vo_dfunc", "vo_ib_ray_jitter", "vo_hydset_u", "vo_hydset_utau",
                 "vo_roughlog_u", "vo_roughlog_utau"]:
        getattr(L, name).restype = d
    for name in ["vo_wf_integrate_F", "vo_wf_find_utau"]:
        getattr(L, name).argtypes = [d, d, d] + ([d] if name == "vo_wf_find_utau" else [])
    L.vo_wf_table.restype = C.POINTER(d)
    L.vo_hydset_u.argtypes = [d, d, d, d]
    L.vo_hydset_utau.argtypes = [d, d, d, d, d]
    L.vo_roughlog_u.argtypes = [d, d, d]
    L.vo_roughlog_utau.argtypes = [d, d, d, d]
    return L


def ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(C.POINTER(C.c_double))
    if a.dtype == np.int32:
        return a.ctypes.data_as(C.POINTER(C.c_int))
    raise TypeError(a.dtype)


def make_cfg(d):
    c = VoCfg()
    for k, v in d.items():
        if k == "bctype":
            for n in range(6):
                c.bctype[n] = int(v[n])
        elif hasattr(c, k):
            setattr(c, k, v)
    return c


class Oracle:
    """Holds a configuration plus metrics and wraps every oracle routine."""

    def __init__(self, cfg: dict, coor: np.ndarray):
        self.cfgd = dict(cfg)
        self.c = make_cfg(cfg)
        self.mx, self.my, self.mz = cfg["mx"], cfg["my"], cfg["mz"]
        self.P = self.mx * self.my * self.mz
        self.g = {}
        for n in GEOM_FIELDS:
            self.g[n] = np.zeros(self.P * (3 if n in VECTOR_GEOM else 1))
        self.geom = VoGeom(*[ptr(self.g[n]) for n in GEOM_FIELDS])
        rc = lib().vo_form_metrics(C.byref(self.c), ptr(coor), C.byref(self.geom))
        if rc:
            raise RuntimeError(f"vo_form_metrics rc={rc}")
        self.nvert = np.zeros(self.P)

    def vec(self):
        return np.zeros(3 * self.P)

    def sca(self):
        return np.zeros(self.P)

    def nvert_sphere(self, cx, cy, cz, r):
        rc = lib().vo_nvert_sphere(C.byref(self.c), C.byref(self.geom), C.c_double(cx), C.c_double(cy),
                                   C.c_double(cz), C.c_double(r), ptr(self.nvert))
        assert rc == 0, rc
        return self.nvert

    def _call(self, name, *args):
        rc = getattr(lib(), name)(C.byref(self.c), *args)
        if isinstance(rc, int) and rc < 0:
            raise RuntimeError(f"{name} rc={rc}")
        return rc

    def contra2cart(self, ucont, ucat):
        self._call("vo_contra2cart", C.byref(self.geom), ptr(self.nvert), ptr(ucont), ptr(ucat))

    def contra2cart_ustar(self, ucont, ucat, ustar):
        self._call("vo_contra2cart_ustar", C.byref(self.geom), ptr(self.nvert), ptr(ucont), ptr(ucat), ptr(ustar))

    def wall_mask(self):
        self._call("vo_wall_mask", ptr(self.nvert))
        return self.nvert

    def eddy_viscosity(self, ucat, cs, nu_t):
        self._call("vo_eddy_viscosity", C.byref(self.geom), ptr(self.nvert), ptr(ucat), C.c_double(cs), ptr(nu_t))

    def eddy_viscosity_cs(self, ucat, cs, nu_t):
        self._call("vo_eddy_viscosity_cs", C.byref(self.geom), ptr(self.nvert), ptr(ucat), ptr(cs), ptr(nu_t))

    def smagorinsky_constant(self, ucat, cs_prev=None, **opt):
        """opt: les, ti, tistart, rstart, max_cs, i_homo, j_homo, k_homo, ik"""
        cs = self.sca() if cs_prev is None else np.array(cs_prev, dtype=np.float64)
        o = les_opt(**opt)
        self._call("vo_smagorinsky_constant", C.byref(self.geom), ptr(self.nvert), ptr(ucat), C.byref(o), ptr(cs))
        return cs

    def dsm_pass1(self, ucat, **opt):
        grad, sabs, uf = np.zeros(9 * self.P), self.sca(), self.vec()
        o = les_opt(**opt)
        self._call("vo_dsm_pass1", C.byref(self.geom), ptr(self.nvert), ptr(ucat), C.byref(o), ptr(grad), ptr(sabs),
                   ptr(uf))
        return grad, sabs, uf

    def dsm_pass2(self, ucat, grad, sabs, uf, **opt):
        LM, MM = self.sca(), self.sca()
        o = les_opt(**opt)
        self._call("vo_dsm_pass2", C.byref(self.geom), ptr(self.nvert), ptr(ucat), C.byref(o), ptr(grad), ptr(sabs),
                   ptr(uf), ptr(LM), ptr(MM))
        return LM, MM

    def dsm_pass3(self, LM, MM, cs_prev=None, **opt):
        cs = self.sca() if cs_prev is None else np.array(cs_prev, dtype=np.float64)
        o = les_opt(**opt)
        self._call("vo_dsm_pass3", C.byref(self.geom), ptr(self.nvert), ptr(LM), ptr(MM), C.byref(o), ptr(cs))
        return cs

    def dsm_smooth(self, cs, **opt):
        o = les_opt(**opt)
        self._call("vo_dsm_smooth", C.byref(self.geom), ptr(self.nvert), C.byref(o), ptr(cs))
        return cs

    def formfunction_snes_f(self, U, U_o, rhs_o, dP, nu_t, f_eul, lucont, ucat, rhs):
        self._call("vo_formfunction_snes_f", C.byref(self.geom), ptr(self.nvert), ptr(U), ptr(U_o), ptr(rhs_o),
                   ptr(dP), ptr(nu_t), None if f_eul is None else ptr(f_eul), ptr(lucont), ptr(ucat), ptr(rhs))

    def rotor_elements(self, coor, cent, itpwidth=6):
        """cent: (3, n) -> box (6, n) int32, dh (3, n)"""
        cent = np.ascontiguousarray(cent, dtype=np.float64)
        n = cent.shape[1]
        box = np.zeros((6, n), dtype=np.int32)
        dh = np.zeros((3, n))
        self._call("vo_rotor_elements", C.byref(self.geom), ptr(coor), C.c_int(n), ptr(cent), C.c_int(itpwidth),
                   ptr(box), ptr(dh))
        return box, dh

    def calc_f_eul(self, coor, cent, F, dA, box, df=10, halfwidth=4.0, width_mode=0, dh_fixed=None):
        cent = np.ascontiguousarray(cent, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        dA = np.ascontiguousarray(dA, dtype=np.float64)
        box = np.ascontiguousarray(box, dtype=np.int32)
        fx = np.ascontiguousarray(dh_fixed if dh_fixed is not None else np.zeros(3), dtype=np.float64)
        f = self.vec()
        self._call("vo_calc_f_eul", C.byref(self.geom), ptr(self.nvert), ptr(coor), C.c_int(cent.shape[1]), ptr(cent),
                   ptr(F), ptr(dA), ptr(box), C.c_int(df), C.c_double(halfwidth), C.c_int(width_mode), ptr(fx),
                   ptr(f))
        return f

    def ib_search(self, body, cap=None):
        """ibm_search_advanced's node list -> structured array (IB_NODE_DTYPE)."""
        from vfsgpu import IB_NODE_DTYPE
        cap = int(cap if cap is not None else max(1, self.P // 8))
        nodes = np.zeros(cap, IB_NODE_DTYPE)
        n = C.c_int(0)
        rc = lib().vo_ib_search(C.byref(self.c), C.byref(self.geom), ptr(self.nvert), C.byref(body.c),
                                nodes.ctypes.data_as(C.c_void_p), C.c_int(cap), C.byref(n))
        if rc:
            raise RuntimeError(f"vo_ib_search rc={rc}")
        return nodes[:n.value].copy()

    def inlet_area(self):
        ka = np.zeros(self.mz)
        m, a = C.c_double(), C.c_double()
        self._call("vo_inlet_area", C.byref(self.geom), ptr(self.nvert), ptr(ka), C.byref(m), C.byref(a))
        return ka, m.value, a.value

    def inflow_flux(self, coor, ucont, ucat, ubcs, inletprofile=1, inlet_flux=-1.0, inlet_area=0.0, uin_plane=None):
        o = VoInflowOpt(int(inletprofile), 0, float(inlet_flux), float(inlet_area))
        f = C.c_double()
        self._call("vo_inflow_flux", C.byref(self.geom), ptr(self.nvert), ptr(coor), C.byref(o),
                   None if uin_plane is None else ptr(uin_plane), ptr(ucont), ptr(ucat), ptr(ubcs), C.byref(f))
        return f.value

    def outflow_flux(self, ucont):
        f = C.c_double()
        self._call("vo_outflow_flux", ptr(self.nvert), ptr(ucont), C.byref(f))
        return f.value

    def outlet_correction(self, flux_in, ucat, ucont, ubcs, scale=True):
        f, a = C.c_double(), C.c_double()
        self._call("vo_outlet_correction", C.byref(self.geom), ptr(self.nvert), C.c_double(flux_in), C.c_int(int(scale)),
                   ptr(ucat), ptr(ucont), ptr(ubcs), C.byref(f), C.byref(a))
        return f.value, a.value

    def ib_averages(self, body, nodes, usum, u1sum, u2sum, N, wallfunction=False, ks=1e-4, roughness=None):
        nodes = np.ascontiguousarray(nodes)
        out = [np.zeros(body.n_elmt) for _ in range(4)]
        o = VoIbOpt(0, _wall_law(wallfunction, roughness), float(ks), float(roughness or 0.0))
        self._call("vo_ib_averages", C.byref(self.geom), C.byref(body.c), nodes.ctypes.data_as(C.c_void_p),
                   C.c_int(len(nodes)), C.byref(o), ptr(usum), ptr(u1sum), ptr(u2sum), C.c_double(N),
                   *[ptr(a) for a in out])
        return out

    def nvert_body(self, body, reset=True, thin=False):
        """ibm_search_advanced's marking of one body into self.nvert."""
        self._call("vo_nvert_body_thin", C.byref(self.geom), ptr(self.nvert), C.byref(body.c),
                   C.c_int(int(bool(reset))), C.c_int(int(bool(thin))))
        return self.nvert

    def ib_interpolate(self, body, nodes, ucont, ucat, p, ustar, moving=False, wallfunction=False, ks=1e-4,
                       roughness=None):
        """ibm_interpolation_advanced; ucont, ucat, p, ustar in/out.
        Returns the element (count, mean ustar^2, mean p)."""
        nodes = np.ascontiguousarray(nodes)
        cnt = np.zeros(body.n_elmt, np.int32)
        sh, pr = np.zeros(body.n_elmt), np.zeros(body.n_elmt)
        o = VoIbOpt(int(bool(moving)), _wall_law(wallfunction, roughness), float(ks), float(roughness or 0.0))
        self._call("vo_ib_interpolate_opt", C.byref(self.geom), ptr(self.nvert), C.byref(body.c),
                   nodes.ctypes.data_as(C.c_void_p), C.c_int(len(nodes)), C.byref(o), ptr(ucont), ptr(ucat), ptr(p),
                   ptr(ustar), ptr(cnt), ptr(sh), ptr(pr))
        return cnt, sh, pr

    def runge_kutta(self, ucont, ucat, nu_t, dP, coor=None, inflow=None, uin_plane=None, ubcs=None, body=None,
                    nodes=None, ib=None, p=None, ustar=None, f_eul=None):
        """vo_runge_kutta (rhs.c:1578-1700): ucont, ucat (and ubcs, p,
        ustar) in/out -> (u0, rhs, flux_in, flux_out).  inflow: a
        VoInflowOpt; ib: a VoIbOpt."""
        u0, rhs = self.vec(), self.vec()
        fi, fo = C.c_double(), C.c_double()
        nodes = None if nodes is None else np.ascontiguousarray(nodes)
        self._call("vo_runge_kutta", C.byref(self.geom), ptr(self.nvert), None if coor is None else ptr(coor),
                   None if inflow is None else C.byref(inflow), None if uin_plane is None else ptr(uin_plane),
                   None if ubcs is None else ptr(ubcs), None if body is None else C.byref(body.c),
                   None if nodes is None else nodes.ctypes.data_as(C.c_void_p),
                   C.c_int(0 if nodes is None else len(nodes)), None if ib is None else C.byref(ib),
                   None if p is None else ptr(p), None if ustar is None else ptr(ustar),
                   None if f_eul is None else ptr(f_eul), ptr(ucont), ptr(ucat), ptr(nu_t), ptr(dP), ptr(u0), ptr(rhs),
                   C.byref(fi), C.byref(fo))
        return u0, rhs, fi.value, fo.value

    def pressure_gradient(self, p, dp, kflux=None, inlet_flux=-1.0):
        """kflux: (mean_k_flux, mean_k_area) from calc_k_flux for the flux
        forcing of a k-periodic context without dpdz_set."""
        if kflux is None:
            self._call("vo_pressure_gradient", C.byref(self.geom), ptr(self.nvert), ptr(p), ptr(dp))
        else:
            self._call("vo_pressure_gradient_kf", C.byref(self.geom), ptr(self.nvert), ptr(p), ptr(dp),
                       C.c_double(kflux[0]), C.c_double(inlet_flux), C.c_double(kflux[1]))

    def calc_k_flux(self, ucont):
        f, a = C.c_double(), C.c_double()
        self._call("vo_calc_k_flux", C.byref(self.geom), ptr(self.nvert), ptr(ucont), C.byref(f), C.byref(a))
        return f.value, a.value

    def ib_bc(self, ucont, ucat):
        self._call("vo_ib_bc", C.byref(self.geom), ptr(self.nvert), ptr(ucont), ptr(ucat))

    def formfunction2(self, ucont, ucat, nu_t, rhs, scale):
        self._call("vo_formfunction2", C.byref(self.geom), ptr(self.nvert), ptr(ucont), ptr(ucat),
                   ptr(nu_t), ptr(rhs), C.c_double(scale))

    def formfunction_snes(self, U, U_o, rhs_o, dP, nu_t, lucont, ucat, rhs):
        self._call("vo_formfunction_snes", C.byref(self.geom), ptr(self.nvert), ptr(U), ptr(U_o),
                   ptr(rhs_o), ptr(dP), ptr(nu_t), ptr(lucont), ptr(ucat), ptr(rhs))

    def rk_stage(self, u0, ucont, ucat, nu_t, dP, alfa, rhs):
        self._call("vo_rk_stage", C.byref(self.geom), ptr(self.nvert), ptr(u0), ptr(ucont), ptr(ucat),
                   ptr(nu_t), ptr(dP), C.c_double(alfa), ptr(rhs))

    def min_dt(self, ucont):
        return lib().vo_min_dt(C.byref(self.c), C.byref(self.geom), ptr(ucont))

    def setup_gid(self):
        self.gid = self.sca()
        self.n = self._call("vo_setup_gid", ptr(self.nvert), ptr(self.gid))
        return self.gid

    def poisson_lhs(self):
        if not hasattr(self, "gid"):
            self.setup_gid()
        n = self.n
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(19 * n + 1, dtype=np.int32)
        val = np.zeros(19 * n + 1)
        nnz = self._call("vo_poisson_lhs", C.byref(self.geom), ptr(self.nvert), ptr(self.gid),
                         ptr(rp), ptr(col), ptr(val))
        self.csr = (rp, col[:nnz].copy(), val[:nnz].copy())
        return self.csr

    def poisson_stencil(self):
        coef = np.zeros(19 * self.P)
        self._call("vo_poisson_stencil", C.byref(self.geom), ptr(self.nvert), ptr(coef))
        return coef

    def poisson_rhs(self, ucont, b):
        self._call("vo_poisson_rhs", C.byref(self.geom), ptr(self.nvert), ptr(ucont), ptr(b))

    def volume_flux(self, ucont):
        f = C.c_double()
        a = C.c_double()
        self._call("vo_volume_flux", C.byref(self.geom), ptr(self.nvert), ptr(ucont), C.byref(f), C.byref(a))
        return f.value, a.value

    def poisson_solve(self, b, phi, rtol=1e-13, maxit=20000):
        if not hasattr(self, "csr"):
            self.poisson_lhs()
        rp, col, val = self.csr
        rel = C.c_double()
        it = self._call("vo_poisson_solve", ptr(self.nvert), ptr(self.gid), ptr(rp), ptr(col), ptr(val),
                        ptr(b), C.c_double(rtol), C.c_int(maxit), ptr(phi), C.byref(rel))
        return it, rel.value

    def csr_apply(self, x, y):
        rp, col, val = self.csr
        self._call("vo_csr_apply", ptr(self.gid), ptr(rp), ptr(col), ptr(val), ptr(x), ptr(y))

    def update_pressure(self, p, phi):
        self._call("vo_update_pressure", ptr(self.nvert), ptr(p), ptr(phi))

    def projection(self, phi, ucont):
        self._call("vo_projection", C.byref(self.geom), ptr(self.nvert), ptr(phi), ptr(ucont))

    def divergence(self, ucont):
        div = self.sca()
        m = lib().vo_divergence(C.byref(self.c), C.byref(self.geom), ptr(self.nvert), ptr(ucont), ptr(div))
        return m, div


# the Cabot wall model of the wall-function walls (wallfunction.c)
WF_NYP = 500000


def dfunc(df, r, n=4.0):
    return lib().vo_dfunc(C.c_int(df), C.c_double(r), C.c_double(n))


def testfilter_ik(val):
    val = np.ascontiguousarray(val, dtype=np.float64)
    assert val.shape == (3, 3, 3)
    return lib().vo_testfilter_ik(ptr(val))


def testfilter(val, w):
    """integrate_testfilter of 3x3x3 arrays indexed [k][j][i] offset + 1."""
    val = np.ascontiguousarray(val, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    assert val.shape == w.shape == (3, 3, 3)
    return lib().vo_testfilter(ptr(val), ptr(w))


def wf_table():
    return np.ctypeslib.as_array(lib().vo_wf_table(), shape=(WF_NYP + 1,)).copy()


def wf_integrate_F(nu, utau, yb):
    return lib().vo_wf_integrate_F(nu, utau, yb)


def wf_find_utau(nu, u, y, guess=0.01):
    return lib().vo_wf_find_utau(nu, u, y, guess)


def wall_function(nu, sc, sb, Uc, n):
    """-> (Ub, ustar) of wall_function with Ua = 0."""
    Uc = np.ascontiguousarray(Uc, dtype=np.float64)
    Ub = np.zeros(3)
    us = C.c_double()
    d = C.c_double
    lib().vo_wall_function(d(nu), d(sc), d(sb), ptr(Uc), ptr(Ub), C.byref(us), d(n[0]), d(n[1]), d(n[2]))
    return Ub, us.value
