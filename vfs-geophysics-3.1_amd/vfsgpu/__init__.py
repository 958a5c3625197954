"""Host-side binding of libvfsgpu.so, the MI355X implementation of the
VFS-Geophysics fractional-step hot path (momentum RHS + pressure projection).

Mirrors the reference's per-step call sequence (solvers.c:562-1520):
``Ctx`` stands for UserCtx; its methods are the C-ABI entry points declared in
include/vfsgpu.h.  Device buffers are plain device pointers owned by
``DeviceArray``.  There is no CPU fallback: if the HIP library is missing or no
GPU is present, the calls fail with an exception.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(_PKG), "csrc")
LIBPATH = os.path.join(CSRC, "libvfsgpu.so")

_lib = None


def lib():
    """Load libvfsgpu.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with __graft_entry__.build()")
        _lib = C.CDLL(LIBPATH)
        _lib.vfs_last_error.restype = C.c_char_p
        _lib.vfs_wl_uniform.restype = C.c_double
        _lib.vfs_wl_uniform.argtypes = [C.c_uint64, C.c_uint64]
    return _lib


class VfsError(RuntimeError):
    def __init__(self, msg, rc=-1):
        super().__init__(msg)
        self.rc = rc


class VfsNotConverged(VfsError):
    """vfs_poisson_solve stopped at maxit above rtol (return code 1); the
    solution array holds the last iterate."""

    def __init__(self, msg, iters, relres):
        super().__init__(msg, 1)
        self.iters, self.relres = iters, relres


def _ck(rc, what):
    if rc != 0:
        raise VfsError(f"{what} failed ({rc}): {lib().vfs_last_error().decode()}", rc)


class Config(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
                ("i_periodic", C.c_int), ("j_periodic", C.c_int), ("k_periodic", C.c_int),
                ("bctype", C.c_int * 6), ("immersed", C.c_int), ("les", C.c_int),
                ("second_order", C.c_int), ("dpdz_set", C.c_int), ("weno", C.c_int), ("inviscid", C.c_int),
                ("skew", C.c_int), ("clark", C.c_int), ("ren", C.c_double), ("dt", C.c_double), ("st", C.c_double),
                ("mean_pressure_gradient", C.c_double)]

    @classmethod
    def from_dict(cls, d):
        c = cls()
        for k, v in d.items():
            if k == "bctype":
                for n in range(6):
                    c.bctype[n] = int(v[n])
            elif k in dict(cls._fields_):
                setattr(c, k, v)
        return c


class LesOptions(C.Structure):
    _fields_ = [("les", C.c_int), ("i_homo_filter", C.c_int), ("j_homo_filter", C.c_int),
                ("k_homo_filter", C.c_int), ("testfilter_ik", C.c_int), ("reserved", C.c_int), ("max_cs", C.c_double)]


class InflowOptions(C.Structure):
    _fields_ = [("inletprofile", C.c_int), ("reserved", C.c_int), ("inlet_flux", C.c_double),
                ("inlet_area", C.c_double)]


class IbOptions(C.Structure):
    _fields_ = [("moving", C.c_int), ("wallfunction", C.c_int), ("ks", C.c_double), ("roughness", C.c_double)]


def _wall_law(wallfunction, roughness):
    """vfs_ib_options.wallfunction: 0 off, 1 wall_function_roughness_a, 2 the
    rough log law of -roughness (ibm.c:3468-3474)."""
    if not wallfunction:
        return 0
    return 1 if roughness is None else 2


class RkBc(C.Structure):
    """vfs_rk_bc: the boundary work of vfs_runge_kutta."""
    _fields_ = [("inflow", C.c_void_p), ("coor", C.c_void_p), ("uin_plane", C.c_void_p), ("ubcs", C.c_void_p),
                ("surf", C.c_void_p), ("nodes", C.c_void_p), ("n_nodes", C.c_int), ("add_body_force", C.c_int),
                ("ib", C.c_void_p), ("p", C.c_void_p), ("ustar", C.c_void_p)]


class IbSurfaceC(C.Structure):
    """vfs_ib_surface: device pointers of a triangulated body."""
    _fields_ = [("n_v", C.c_int), ("n_elmt", C.c_int),
                ("x_bp", C.c_void_p), ("y_bp", C.c_void_p), ("z_bp", C.c_void_p),
                ("nv1", C.c_void_p), ("nv2", C.c_void_p), ("nv3", C.c_void_p),
                ("nf_x", C.c_void_p), ("nf_y", C.c_void_p), ("nf_z", C.c_void_p), ("u", C.c_void_p)]


# vfs_ib_node as a numpy record (the C struct's layout: 14 ints, 11 doubles)
IB_NODE_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("k", "<i4"), ("cell", "<i4"), ("imode", "<i4"),
                          ("i1", "<i4"), ("j1", "<i4"), ("k1", "<i4"), ("i2", "<i4"), ("j2", "<i4"), ("k2", "<i4"),
                          ("i3", "<i4"), ("j3", "<i4"), ("k3", "<i4"),
                          ("cr1", "<f8"), ("cr2", "<f8"), ("cr3", "<f8"), ("cs1", "<f8"), ("cs2", "<f8"),
                          ("cs3", "<f8"), ("d_s", "<f8"), ("d_i", "<f8"), ("pmin", "<f8", (3,))])
assert IB_NODE_DTYPE.itemsize == 144


class DeviceArray:
    """A zero-initialised device allocation of n float64 (or float32) values."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _ck(lib().vfs_device_malloc(C.c_size_t(self.nbytes), C.byref(p)), "vfs_device_malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _ck(lib().vfs_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes)), "htod")

    def host(self):
        out = np.empty(self.n, self.dtype)
        _ck(lib().vfs_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes)), "dtoh")
        return out

    def copy_from(self, other):
        _ck(lib().vfs_memcpy_dtod(self.ptr, other.ptr, C.c_size_t(self.nbytes)), "dtod")

    def free(self):
        if self.ptr:
            lib().vfs_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(x):
    if x is None:
        return None
    return x.ptr


METRIC_NAMES = ["csi", "eta", "zet", "aj", "icsi", "ieta", "izet", "iaj",
                "jcsi", "jeta", "jzet", "jaj", "kcsi", "keta", "kzet", "kaj", "cent"]
SCALAR_METRICS = {"aj", "iaj", "jaj", "kaj"}


class Ctx:
    """One grid block: the reference's UserCtx with metrics and nvert on the GPU."""

    def __init__(self, cfg: dict):
        self.cfgd = dict(cfg)
        self.cfg = Config.from_dict(cfg)
        self.mx, self.my, self.mz = cfg["mx"], cfg["my"], cfg["mz"]
        self.P = self.mx * self.my * self.mz
        h = C.c_void_p()
        _ck(lib().vfs_ctx_create(C.byref(self.cfg), C.byref(h)), "vfs_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().vfs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def vec(self):
        return DeviceArray(3 * self.P)

    def sca(self):
        return DeviceArray(self.P)

    def sync(self):
        _ck(lib().vfs_synchronize(self.h), "vfs_synchronize")

    # geometry
    def build_metrics(self, coor):
        _ck(lib().vfs_build_metrics(self.h, coor.ptr), "vfs_build_metrics")

    def metric(self, name):
        w = METRIC_NAMES.index(name)
        out = DeviceArray(self.P * (1 if name in SCALAR_METRICS else 3))
        _ck(lib().vfs_get_metric(self.h, C.c_int(w), out.ptr), "vfs_get_metric")
        return out

    def set_nvert(self, nvert):
        _ck(lib().vfs_set_nvert(self.h, nvert.ptr), "vfs_set_nvert")

    def nvert(self):
        out = self.sca()
        _ck(lib().vfs_get_nvert(self.h, out.ptr), "vfs_get_nvert")
        return out

    def nvert_sphere(self, cx, cy, cz, r):
        out = self.sca()
        _ck(lib().vfs_nvert_sphere(self.h, C.c_double(cx), C.c_double(cy), C.c_double(cz), C.c_double(r), out.ptr),
            "vfs_nvert_sphere")
        return out

    # momentum
    def contra2cart(self, ucont, ucat):
        _ck(lib().vfs_contra2cart(self.h, ucont.ptr, ucat.ptr), "vfs_contra2cart")

    def eddy_viscosity(self, ucat, cs, nu_t):
        _ck(lib().vfs_eddy_viscosity(self.h, ucat.ptr, C.c_double(cs), nu_t.ptr), "vfs_eddy_viscosity")

    def eddy_viscosity_cs(self, ucat, cs, nu_t):
        """nu_t from a per-cell Cs field (DeviceArray of P doubles)."""
        _ck(lib().vfs_eddy_viscosity_cs(self.h, ucat.ptr, cs.ptr, nu_t.ptr), "vfs_eddy_viscosity_cs")

    def rotor_elements(self, coor, cent, itpwidth=6):
        """cent: DeviceArray of 3n doubles (x|y|z blocks) -> (box, dh) as
        DeviceArrays (box holds 6n int32 values)."""
        n = cent.n // 3
        box = DeviceArray(6 * n, np.int32)
        dh = DeviceArray(3 * n)
        _ck(lib().vfs_rotor_elements(self.h, coor.ptr, int(n), cent.ptr, int(itpwidth), box.ptr, dh.ptr),
            "vfs_rotor_elements")
        return box, dh

    def calc_f_eul(self, coor, cent, f_lagr, dA, box, f_eul, deltafunc=10, halfwidth=4.0, width_mode=0,
                   dh_fixed=None):
        fx = None if dh_fixed is None else (C.c_double * 3)(*dh_fixed)
        _ck(lib().vfs_calc_f_eul(self.h, coor.ptr, int(cent.n // 3), cent.ptr, f_lagr.ptr, dA.ptr, box.ptr,
                                 int(deltafunc), C.c_double(halfwidth), int(width_mode), fx, f_eul.ptr),
            "vfs_calc_f_eul")

    def inlet_area(self):
        """Calc_Inlet_Area (bcs.c:216-328) -> (k_area DeviceArray of mz,
        mean_k_area, inlet_area)."""
        ka = DeviceArray(self.mz)
        m, a = C.c_double(), C.c_double()
        _ck(lib().vfs_inlet_area(self.h, ka.ptr, C.byref(m), C.byref(a)), "vfs_inlet_area")
        return ka, m.value, a.value

    def inflow_flux(self, coor, ucont, ucat, ubcs, inletprofile=1, inlet_flux=-1.0, inlet_area=0.0, uin_plane=None):
        """InflowFlux (bcs.c:329-706) on the k = 0 plane -> the inflow."""
        o = InflowOptions(int(inletprofile), 0, float(inlet_flux), float(inlet_area))
        f = C.c_double()
        _ck(lib().vfs_inflow_flux(self.h, C.byref(o), coor.ptr, None if uin_plane is None else uin_plane.ptr,
                                  ucont.ptr, ucat.ptr, ubcs.ptr, C.byref(f)), "vfs_inflow_flux")
        return f.value

    def outflow_flux(self, ucont):
        f = C.c_double()
        _ck(lib().vfs_outflow_flux(self.h, ucont.ptr, C.byref(f)), "vfs_outflow_flux")
        return f.value

    def outlet_correction(self, flux_in, ucat, ucont, ubcs, scale=True):
        """FormBCS's outlet branch (bcs.c:2285-2418) -> (flux_out, area)."""
        f, a = C.c_double(), C.c_double()
        _ck(lib().vfs_outlet_correction(self.h, C.c_double(flux_in), int(bool(scale)), ucat.ptr, ucont.ptr, ubcs.ptr,
                                        C.byref(f), C.byref(a)), "vfs_outlet_correction")
        return f.value, a.value

    def nvert_body(self, surface, reset=True, thin=False):
        """ibm_search_advanced's marking (ibm.c:110-360) of one triangulated
        body into the context's mask (thin: -thin) -> the mask as a
        DeviceArray."""
        out = self.sca()
        _ck(lib().vfs_nvert_body_thin(self.h, C.byref(surface.c), int(bool(reset)), int(bool(thin)), out.ptr),
            "vfs_nvert_body_thin")
        return out

    def ib_search(self, surface, cap=None):
        """ibm_search_advanced's node list (ibm.c:366-497) for the current
        mask -> (nodes DeviceArray of 144-byte vfs_ib_node records, n)."""
        cap = int(cap if cap is not None else max(1, self.P // 8))
        nodes = DeviceArray(cap * IB_NODE_DTYPE.itemsize, np.uint8)
        n = C.c_int(0)
        _ck(lib().vfs_ib_search(self.h, C.byref(surface.c), nodes.ptr, C.c_int(cap), C.byref(n)), "vfs_ib_search")
        return nodes, n.value

    def ib_search_noalloc(self, surface, nodes, cap):
        """vfs_ib_search_noalloc into a caller's node list (a DeviceArray of
        cap 144-byte records) -> the node count."""
        n = C.c_int(0)
        _ck(lib().vfs_ib_search_noalloc(self.h, C.byref(surface.c), nodes.ptr, C.c_int(int(cap)), C.byref(n)),
            "vfs_ib_search_noalloc")
        return n.value

    def ib_interpolate(self, surface, nodes, n, ucont, ucat, p, ustar, stats=True, moving=False, wallfunction=False,
                       ks=1e-4, roughness=None):
        """ibm_interpolation_advanced (ibm.c:3294-3698): ucont, ucat, p, ustar
        in/out.  Returns the element (count, mean ustar^2, mean p) as
        DeviceArrays, or None with stats=False.  With wallfunction, a
        roughness (-roughness, set for body 0 only in the reference) selects
        the rough log law instead of wall_function_roughness_a."""
        ne = surface.n_elmt
        st = (DeviceArray(ne, np.int32), DeviceArray(ne), DeviceArray(ne)) if stats else (None, None, None)
        o = IbOptions(int(bool(moving)), _wall_law(wallfunction, roughness), float(ks),
                      float(roughness or 0.0))
        _ck(lib().vfs_ib_interpolate_opt(self.h, C.byref(surface.c), nodes.ptr, C.c_int(int(n)), C.byref(o), ucont.ptr,
                                         ucat.ptr, p.ptr, ustar.ptr, *[None if a is None else a.ptr for a in st]),
            "vfs_ib_interpolate_opt")
        return st if stats else None

    def ib_averages(self, surface, nodes, n, usum, u1sum, u2sum, N, wallfunction=False, ks=1e-4, roughness=None):
        """The averaging statistics of ibm_interpolation_advanced -> the
        element means (ustar_avg^2, and the three contravariant Reynolds
        stresses) as DeviceArrays."""
        ne = surface.n_elmt
        out = [DeviceArray(ne) for _ in range(4)]
        o = IbOptions(0, _wall_law(wallfunction, roughness), float(ks), float(roughness or 0.0))
        _ck(lib().vfs_ib_averages(self.h, C.byref(surface.c), nodes.ptr, C.c_int(int(n)), C.byref(o), usum.ptr,
                                  u1sum.ptr, u2sum.ptr, C.c_double(N), *[a.ptr for a in out]), "vfs_ib_averages")
        return out

    def set_body_force(self, f_eul):
        _ck(lib().vfs_set_body_force(self.h, None if f_eul is None else f_eul.ptr), "vfs_set_body_force")

    def smagorinsky_constant(self, ucat, cs, les=2, ti=2, tistart=0, rstart=0, max_cs=0.5, i_homo=0, j_homo=0,
                             k_homo=0, ik=0):
        """Compute_Smagorinsky_Constant_1 (les.c:67-1133) into cs: -les 1
        (0.01), 2 (dynamic procedure) or 3 (relaxed; cs also holds the
        previous Cs); 0 while ti < 2 on a fresh start."""
        o = LesOptions(int(les), int(i_homo), int(j_homo), int(k_homo), int(ik), 0, float(max_cs))
        _ck(lib().vfs_smagorinsky_constant(self.h, ucat.ptr, C.byref(o), int(ti), int(tistart), int(rstart), cs.ptr),
            "vfs_smagorinsky_constant")

    def set_inlet_flux(self, inlet_flux):
        """-flux: the target k flux of the forcing of a k-periodic context
        without dpdz_set (the reference's default -1)."""
        _ck(lib().vfs_set_inlet_flux(self.h, C.c_double(inlet_flux)), "vfs_set_inlet_flux")

    def calc_k_flux(self, ucont, read=True):
        """Calc_k_Flux (rhs2.c:1562-1645) -> (mean_k_flux, mean_k_area), kept
        on the device for Pressure_Gradient (None, None with read=False)."""
        f, a = C.c_double(), C.c_double()
        _ck(lib().vfs_calc_k_flux(self.h, ucont.ptr, C.byref(f) if read else None, C.byref(a) if read else None),
            "vfs_calc_k_flux")
        return (f.value, a.value) if read else (None, None)

    def pressure_gradient(self, p, dp):
        _ck(lib().vfs_pressure_gradient(self.h, p.ptr, dp.ptr), "vfs_pressure_gradient")

    def ib_bc(self, ucont, ucat):
        _ck(lib().vfs_ib_bc(self.h, ucont.ptr, ucat.ptr), "vfs_ib_bc")

    def formfunction2(self, ucont, ucat, nu_t, rhs, scale):
        _ck(lib().vfs_formfunction2(self.h, ucont.ptr, ucat.ptr, _p(nu_t), rhs.ptr, C.c_double(scale)),
            "vfs_formfunction2")

    def formfunction_snes(self, U, U_o, rhs_o, dP, nu_t, lucont, ucat, rhs):
        _ck(lib().vfs_formfunction_snes(self.h, U.ptr, U_o.ptr, rhs_o.ptr, dP.ptr, _p(nu_t), lucont.ptr, ucat.ptr,
                                        rhs.ptr), "vfs_formfunction_snes")

    def rk_stage(self, u0, ucont, ucat, nu_t, dP, alfa, rhs):
        _ck(lib().vfs_rk_stage(self.h, u0.ptr, ucont.ptr, ucat.ptr, _p(nu_t), dP.ptr, C.c_double(alfa), rhs.ptr),
            "vfs_rk_stage")

    def runge_kutta(self, ucont, ucat, nu_t, dP, u0, rhs, inflow=None, coor=None, uin_plane=None, ubcs=None,
                    surface=None, nodes=None, n_nodes=0, ib=None, p=None, ustar=None, add_body_force=False):
        """RungeKutta (rhs.c:1578-1700) with its boundary work -> (flux_in,
        flux_out).  inflow: an InflowOptions for an inlet at k = 0; ib: an
        IbOptions (None: the defaults)."""
        keep = []

        def ref(x):
            if x is None:
                return None
            keep.append(x)
            return C.cast(C.byref(x), C.c_void_p)

        bc = RkBc(ref(inflow), _p(coor), _p(uin_plane), _p(ubcs), ref(surface.c) if surface is not None else None,
                  _p(nodes), int(n_nodes), int(bool(add_body_force)), ref(ib), _p(p), _p(ustar))
        fi, fo = C.c_double(), C.c_double()
        _ck(lib().vfs_runge_kutta(self.h, C.byref(bc), ucont.ptr, ucat.ptr, _p(nu_t), dP.ptr, u0.ptr, rhs.ptr,
                                  C.byref(fi), C.byref(fo)), "vfs_runge_kutta")
        return fi.value, fo.value

    def min_dt(self, ucont):
        v = C.c_double()
        _ck(lib().vfs_min_dt(self.h, ucont.ptr, C.byref(v)), "vfs_min_dt")
        return v.value

    def rhs_jvp(self, U, ucat, nu_t, w, jw):
        _ck(lib().vfs_rhs_jvp(self.h, U.ptr, ucat.ptr, _p(nu_t), w.ptr, jw.ptr), "vfs_rhs_jvp")

    def rhs_vjp(self, U, ucat, nu_t, v, jtv):
        _ck(lib().vfs_rhs_vjp(self.h, U.ptr, ucat.ptr, _p(nu_t), v.ptr, jtv.ptr), "vfs_rhs_vjp")

    def implicit_matrixfree(self, U, U_o, rhs_o, dP, nu_t, ucat, rtol=1e-4, max_newton=50, restart=30,
                            max_krylov=1000):
        """Implicit_MatrixFree: returns (newton_its, krylov_its, final ||F||)."""
        ni, ki, f = C.c_int(), C.c_int(), C.c_double()
        _ck(lib().vfs_implicit_matrixfree(self.h, U.ptr, U_o.ptr, rhs_o.ptr, dP.ptr, _p(nu_t), ucat.ptr,
                                          C.c_double(rtol), C.c_int(max_newton), C.c_int(restart),
                                          C.c_int(max_krylov), C.byref(ni), C.byref(ki), C.byref(f)),
            "vfs_implicit_matrixfree")
        return ni.value, ki.value, f.value

    # pressure
    def poisson_setup(self):
        n = C.c_int()
        _ck(lib().vfs_poisson_setup(self.h, C.byref(n)), "vfs_poisson_setup")
        return n.value

    def poisson_gid(self):
        out = self.sca()
        _ck(lib().vfs_poisson_gid(self.h, out.ptr), "vfs_poisson_gid")
        return out

    def poisson_coefficients(self):
        out = DeviceArray(19 * self.P)
        _ck(lib().vfs_poisson_coefficients(self.h, out.ptr), "vfs_poisson_coefficients")
        return out

    def poisson_apply(self, x, y):
        _ck(lib().vfs_poisson_apply(self.h, x.ptr, y.ptr), "vfs_poisson_apply")

    def volume_flux(self, ucont):
        f, a = C.c_double(), C.c_double()
        _ck(lib().vfs_volume_flux(self.h, ucont.ptr, C.byref(f), C.byref(a)), "vfs_volume_flux")
        return f.value, a.value

    def poisson_rhs(self, ucont, b):
        _ck(lib().vfs_poisson_rhs(self.h, ucont.ptr, b.ptr), "vfs_poisson_rhs")

    def poisson_solve(self, b, phi, rtol=1e-8, precision=0, maxit=500):
        """-> (iterations, relres); raises VfsNotConverged when maxit is
        reached above rtol and VfsError on a non-finite residual."""
        it, rel = C.c_int(), C.c_double()
        rc = lib().vfs_poisson_solve(self.h, b.ptr, phi.ptr, C.c_double(rtol), C.c_int(precision), C.c_int(maxit),
                                     C.byref(it), C.byref(rel))
        if rc == 1:
            raise VfsNotConverged(f"vfs_poisson_solve: {lib().vfs_last_error().decode()}", it.value, rel.value)
        _ck(rc, "vfs_poisson_solve")
        return it.value, rel.value

    def update_pressure(self, p, phi):
        _ck(lib().vfs_update_pressure(self.h, p.ptr, phi.ptr), "vfs_update_pressure")

    def projection(self, phi, ucont):
        _ck(lib().vfs_projection(self.h, phi.ptr, ucont.ptr), "vfs_projection")

    def divergence(self, ucont, div=None):
        m = C.c_double()
        _ck(lib().vfs_divergence(self.h, ucont.ptr, _p(div), C.byref(m)), "vfs_divergence")
        return m.value

    def ustar(self):
        """Friction velocity of the wall-function cells (user->lUstar)."""
        out = self.sca()
        _ck(lib().vfs_get_ustar(self.h, out.ptr), "vfs_get_ustar")
        return out

    def restart_write(self, directory, ti, block, ucont, ucat, p, nvert):
        """Ucont_P_Binary_Output (main.c:799-847): the vfield/ufield/pfield/
        nvfield restart files of block `block` at step `ti`."""
        _ck(lib().vfs_restart_write(self.h, str(directory).encode(), int(ti), int(block), ucont.ptr, ucat.ptr, p.ptr,
                                    nvert.ptr), "vfs_restart_write")

    def restart_read(self, directory, ti, block, ucont, ucat, p, nvert_o):
        """Ucont_P_Binary_Input (main.c:394-420); nvfield goes to nvert_o."""
        _ck(lib().vfs_restart_read(self.h, str(directory).encode(), int(ti), int(block), ucont.ptr, ucat.ptr, p.ptr,
                                   nvert_o.ptr), "vfs_restart_read")


# ---------------------------------------------------------------------------
# synthetic workloads (host generators compiled into the library)
# ---------------------------------------------------------------------------
def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# the reference's file formats (host I/O in the library)
# ---------------------------------------------------------------------------
GRID_TEXT, GRID_BINARY, GRID_XYZ = 0, 1, 2


def read_grid(path, fmt=GRID_TEXT, cl=1.0, L_dim=1.0):
    """Grid file of MG_Initial (init.c:316-436) -> ((IM, JM, KM), coor) with
    coor in the layout Ctx.build_metrics takes ((IM+1)(JM+1)(KM+1) Cmpnts)."""
    dims = (C.c_int * 3)()
    args = (str(path).encode(), int(fmt), C.c_double(cl), C.c_double(L_dim), dims)
    _ck(lib().vfs_read_grid(*args, None), "vfs_read_grid")
    IM, JM, KM = dims[0], dims[1], dims[2]
    coor = np.zeros(3 * (IM + 1) * (JM + 1) * (KM + 1))
    _ck(lib().vfs_read_grid(*args, _dp(coor)), "vfs_read_grid")
    return (IM, JM, KM), coor


def write_petsc_vec(path, a):
    """A host array as a PETSc binary vector file."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    _ck(lib().vfs_vec_write_petsc(str(path).encode(), _dp(a), C.c_longlong(a.size)), "vfs_vec_write_petsc")


def read_petsc_vec(path, n):
    """A PETSc binary vector file of exactly n values."""
    a = np.zeros(int(n))
    _ck(lib().vfs_vec_read_petsc(str(path).encode(), _dp(a), C.c_longlong(a.size)), "vfs_vec_read_petsc")
    return a


def restart_path(directory, stem, ti, block):
    buf = C.create_string_buffer(1024)
    _ck(lib().vfs_restart_path(buf, C.c_size_t(1024), str(directory).encode(), stem.encode(), int(ti), int(block)),
        "vfs_restart_path")
    return buf.value.decode()


def channel_grid(IM, JM, KM, Lx, Ly, Lz, beta=0.0, amp=0.0):
    mx, my, mz = IM + 1, JM + 1, KM + 1
    coor = np.zeros(3 * mx * my * mz)
    _ck(lib().vfs_wl_channel_grid(IM, JM, KM, C.c_double(Lx), C.c_double(Ly), C.c_double(Lz), C.c_double(beta),
                                  C.c_double(amp), _dp(coor)), "vfs_wl_channel_grid")
    return coor


def poiseuille(cfg, Ly, noise, seed, cent, nvert, icsi, jeta, kzet):
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    per = (C.c_int * 3)(cfg["i_periodic"], cfg["j_periodic"], cfg["k_periodic"])
    ucat = np.zeros(3 * mx * my * mz)
    ucont = np.zeros(3 * mx * my * mz)
    _ck(lib().vfs_wl_poiseuille(mx, my, mz, per, C.c_double(Ly), C.c_double(noise), C.c_uint64(seed), _dp(cent),
                                _dp(nvert), _dp(icsi), _dp(jeta), _dp(kzet), _dp(ucat), _dp(ucont)), "poiseuille")
    return ucat, ucont


def loglaw_k(cfg, Ly, y0, noise, seed, cent, nvert, icsi, jeta, kzet):
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    per = (C.c_int * 3)(cfg["i_periodic"], cfg["j_periodic"], cfg["k_periodic"])
    ucat = np.zeros(3 * mx * my * mz)
    ucont = np.zeros(3 * mx * my * mz)
    _ck(lib().vfs_wl_loglaw_k(mx, my, mz, per, C.c_double(Ly), C.c_double(y0), C.c_double(noise), C.c_uint64(seed),
                              _dp(cent), _dp(nvert), _dp(icsi), _dp(jeta), _dp(kzet), _dp(ucat), _dp(ucont)),
        "loglaw_k")
    return ucat, ucont


def turbulent(cfg, Lx, Ly, Lz, nmodes, nmax, urms, seed, cent, nvert, icsi, jeta, kzet):
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    per = (C.c_int * 3)(cfg["i_periodic"], cfg["j_periodic"], cfg["k_periodic"])
    ucat = np.zeros(3 * mx * my * mz)
    ucont = np.zeros(3 * mx * my * mz)
    _ck(lib().vfs_wl_turbulent(mx, my, mz, per, C.c_double(Lx), C.c_double(Ly), C.c_double(Lz), nmodes, nmax,
                               C.c_double(urms), C.c_uint64(seed), _dp(cent), _dp(nvert), _dp(icsi), _dp(jeta),
                               _dp(kzet), _dp(ucat), _dp(ucont)), "turbulent")
    return ucat, ucont


def random_scalar(cfg, amp, seed, nvert=None, aj=None):
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    out = np.zeros(mx * my * mz)
    _ck(lib().vfs_wl_random_scalar(mx, my, mz, C.c_double(amp), C.c_uint64(seed),
                                   _dp(nvert) if nvert is not None else None,
                                   _dp(aj) if aj is not None else None, _dp(out)), "random_scalar")
    return out


def random_vector(cfg, amp, seed):
    mx, my, mz = cfg["mx"], cfg["my"], cfg["mz"]
    out = np.zeros(3 * mx * my * mz)
    _ck(lib().vfs_wl_random_vector(mx, my, mz, C.c_double(amp), C.c_uint64(seed), _dp(out)), "random_vector")
    return out


def uniform(seed, counter):
    return lib().vfs_wl_uniform(seed, counter)


class IbSurface:
    """A triangulated body on the device (the reference's IBMNodes): vertices
    (n_v, 3), elements (n_elmt, 3) int, unit normals (n_elmt, 3) pointing
    into the fluid, optional vertex velocities (n_v, 3)."""

    def __init__(self, verts, elems, normals, u=None):
        verts = np.asarray(verts, np.float64)
        elems = np.asarray(elems, np.int32)
        normals = np.asarray(normals, np.float64)
        self.n_v, self.n_elmt = len(verts), len(elems)
        self._a = [DeviceArray.from_host(np.ascontiguousarray(verts[:, q])) for q in range(3)]
        self._a += [DeviceArray.from_host(np.ascontiguousarray(elems[:, q])) for q in range(3)]
        self._a += [DeviceArray.from_host(np.ascontiguousarray(normals[:, q])) for q in range(3)]
        self.u = None if u is None else DeviceArray.from_host(np.ascontiguousarray(u, np.float64).ravel())
        self.c = IbSurfaceC(self.n_v, self.n_elmt, *[a.ptr for a in self._a], None if self.u is None else self.u.ptr)


def ray_jitter(seed):
    """randomdirection's jitter for a seed (the C library's first rand())."""
    v = C.c_double()
    _ck(lib().vfs_ib_ray_jitter(int(seed), C.byref(v)), "vfs_ib_ray_jitter")
    return v.value


def _wl_uniform_array(seed, n):
    """vfs_wl_uniform(seed, c) for c = 0..n-1 (the library's counter RNG)."""
    f = lib().vfs_wl_uniform
    return np.array([f(int(seed), c) for c in range(int(n))])
