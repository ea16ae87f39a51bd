"""ctypes binding of libnebulasem_amd.so (the C-ABI in include/nebulasem_amd.h).

The library must have been built (nebulasem_amd/build.py, or
__graft_entry__.build()); nothing here falls back to Python or to the CPU.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NEBULASEM_AMD_LIB selects another build of the same library (the host
# sanitizer build, nebulasem_amd/build.py: build_sanitized)
LIB_PATH = os.environ.get("NEBULASEM_AMD_LIB", os.path.join(_HERE, "libnebulasem_amd.so"))

_lib = None


class NsemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nsem status {code}: {msg}")
        self.code = code


class MeshInfo(C.Structure):
    _fields_ = [("npx", C.c_uint32), ("npy", C.c_uint32), ("npz", C.c_uint32),
                ("np", C.c_uint32), ("npf", C.c_uint32), ("nbcs", C.c_uint32),
                ("ncells", C.c_uint32), ("nfacets", C.c_uint32), ("nvertices", C.c_uint32),
                ("nbcs_field", C.c_uint64), ("nall_field", C.c_uint64),
                ("nboundaries", C.c_uint32)]


def lib():
    """Load the native library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with nebulasem_amd/build.py")
        _lib = C.CDLL(LIB_PATH)
        _lib.nsem_last_error.restype = C.c_char_p
        _lib.nsem_version.restype = C.c_char_p
    return _lib


def check(code):
    if code != 0:
        raise NsemError(code, lib().nsem_last_error().decode())


def _i32x3(v):
    return (C.c_int32 * 3)(*v)


class Mesh:
    """A loaded mesh with its node geometry (host side)."""

    def __init__(self, handle):
        self._h = handle
        info = MeshInfo()
        check(lib().nsem_mesh_get_info(self._h, C.byref(info)))
        self.info = info
        self.boundaries = []
        buf = C.create_string_buffer(256)
        for i in range(info.nboundaries):
            check(lib().nsem_mesh_boundary_name(self._h, i, buf, 256))
            self.boundaries.append(buf.value.decode())

    @property
    def handle(self):
        return self._h

    @classmethod
    def load(cls, grid_path, order=(0, 0, 0), nonortho=3):
        h = C.c_void_p()
        check(lib().nsem_mesh_load(grid_path.encode(), _i32x3(order), nonortho, C.byref(h)))
        return cls(h)

    @classmethod
    def from_spec(cls, spec_path, order=(0, 0, 0), nonortho=3):
        h = C.c_void_p()
        check(lib().nsem_mesh_from_spec(spec_path.encode(), _i32x3(order), nonortho, C.byref(h)))
        return cls(h)

    @classmethod
    def box(cls, corners, ndiv, patches=(), order=(0, 0, 0), nonortho=3, spacing=None,
            ratio=None):
        """One hex block.  corners: (8,3) in the reference's block corner order;
        patches: sequence of (name, [(x,y,z), ...])."""
        corners = np.ascontiguousarray(corners, dtype=np.float64).reshape(24)
        names = (C.c_char_p * max(1, len(patches)))(*[p[0].encode() for p in patches])
        npts = (C.c_int32 * max(1, len(patches)))(*[len(p[1]) for p in patches])
        pts = np.ascontiguousarray([q for p in patches for q in p[1]], dtype=np.float64).reshape(-1)
        if pts.size == 0:
            pts = np.zeros(3)
        sp = (C.c_int32 * 12)(*(spacing if spacing is not None else [0] * 12))
        ra = (C.c_double * 12)(*(ratio if ratio is not None else [1.0] * 12))
        h = C.c_void_p()
        check(lib().nsem_mesh_box(corners.ctypes.data_as(C.POINTER(C.c_double)),
                                  (C.c_uint32 * 3)(*ndiv), sp, ra, len(patches), names, npts,
                                  pts.ctypes.data_as(C.POINTER(C.c_double)), _i32x3(order),
                                  nonortho, C.byref(h)))
        return cls(h)

    def array(self, name):
        count = C.c_uint64()
        kind = C.c_int32()
        check(lib().nsem_mesh_array_size(self._h, name.encode(), C.byref(count), C.byref(kind)))
        out = np.empty(count.value, dtype=np.uint32 if kind.value == 0 else np.float64)
        check(lib().nsem_mesh_get_array(self._h, name.encode(), out.ctypes.data_as(C.c_void_p),
                                        C.c_uint64(count.value)))
        return out

    def write_grid(self, path):
        check(lib().nsem_mesh_write_grid(self._h, path.encode()))

    def close(self):
        if self._h:
            lib().nsem_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_copy(nbytes, reps=10):
    """Seconds per launch of a plain device copy of nbytes (nsem_stream_copy)."""
    sec = C.c_double()
    check(lib().nsem_stream_copy(C.c_uint64(nbytes), reps, C.byref(sec)))
    return sec.value


def grid_generate(spec_path, grid_path):
    check(lib().nsem_grid_generate(spec_path.encode(), grid_path.encode()))


def read_field(mesh, path, ncomp, gravity=(0.0, 0.0, 0.0)):
    """Read a reference field file: (values[nall, ncomp], bc_kind[nb], bc_value[nb, ncomp])."""
    n = mesh.info.nall_field
    nb = mesh.info.nboundaries
    vals = np.empty(n * ncomp)
    kinds = np.empty(max(nb, 1), dtype=np.int32)
    bv = np.empty(max(nb, 1) * ncomp)
    check(lib().nsem_field_read(mesh.handle, path.encode(), ncomp, (C.c_double * 3)(*gravity),
                                vals.ctypes.data_as(C.c_void_p), kinds.ctypes.data_as(C.c_void_p),
                                bv.ctypes.data_as(C.c_void_p)))
    return vals.reshape(n, ncomp), kinds[:nb], bv[:nb * ncomp].reshape(nb, ncomp)


def read_field_bc(mesh, path, ncomp, gravity=(0.0, 0.0, 0.0)):
    """read_field with ROBIN patches (nsem_field_read_bc): (values[nall, ncomp],
    bc_kind[nb], bc_value[nb, ncomp], bc_shape[nb], bc_tvalue[nb, ncomp])."""
    n = mesh.info.nall_field
    nb = mesh.info.nboundaries
    vals = np.empty(n * ncomp)
    kinds = np.empty(max(nb, 1), dtype=np.int32)
    bv = np.empty(max(nb, 1) * ncomp)
    sh = np.empty(max(nb, 1))
    tv = np.empty(max(nb, 1) * ncomp)
    check(lib().nsem_field_read_bc(mesh.handle, path.encode(), ncomp, (C.c_double * 3)(*gravity),
                                   vals.ctypes.data_as(C.c_void_p), kinds.ctypes.data_as(C.c_void_p),
                                   bv.ctypes.data_as(C.c_void_p), sh.ctypes.data_as(C.c_void_p),
                                   tv.ctypes.data_as(C.c_void_p)))
    return (vals.reshape(n, ncomp), kinds[:nb], bv[:nb * ncomp].reshape(nb, ncomp), sh[:nb],
            tv[:nb * ncomp].reshape(nb, ncomp))


def read_amr_tree(path):
    """amrTree_<step>.bin/.txt: (n, 4) uint32 nodes (id, nchildren, cid, level)."""
    n = C.c_uint64()
    check(lib().nsem_amr_tree_read(path.encode(), None, C.c_uint64(0), C.byref(n)))
    out = np.empty((n.value, 4), dtype=np.uint32)
    check(lib().nsem_amr_tree_read(path.encode(), out.ctypes.data_as(C.c_void_p), C.c_uint64(n.value),
                                   C.byref(n)))
    return out


def write_amr_tree(path, nodes):
    nodes = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 4)
    check(lib().nsem_amr_tree_write(path.encode(), nodes.ctypes.data_as(C.c_void_p), C.c_uint64(nodes.shape[0])))


def write_field(mesh, path, values, bc_kind, bc_value):
    """Write a reference field file (nsem_field_write): values (nbcs_field, ncomp)
    interior values; bc_kind (nb,), bc_value (nb, ncomp) per mesh patch."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    ncomp = 1 if v.ndim == 1 else v.shape[1]
    k = np.ascontiguousarray(bc_kind, dtype=np.int32).reshape(-1)
    bv = np.ascontiguousarray(bc_value, dtype=np.float64).reshape(-1)
    check(lib().nsem_field_write(mesh.handle, path.encode(), ncomp, v.ctypes.data_as(C.c_void_p),
                                 k.ctypes.data_as(C.c_void_p), bv.ctypes.data_as(C.c_void_p)))


def write_field_bc(mesh, path, values, bc_kind, bc_value, bc_shape, bc_tvalue):
    """write_field with ROBIN patches (nsem_field_write_bc): bc_shape (nb,),
    bc_tvalue (nb, ncomp)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    ncomp = 1 if v.ndim == 1 else v.shape[1]
    k = np.ascontiguousarray(bc_kind, dtype=np.int32).reshape(-1)
    bv = np.ascontiguousarray(bc_value, dtype=np.float64).reshape(-1)
    sh = np.ascontiguousarray(bc_shape, dtype=np.float64).reshape(-1)
    tv = np.ascontiguousarray(bc_tvalue, dtype=np.float64).reshape(-1)
    check(lib().nsem_field_write_bc(mesh.handle, path.encode(), ncomp, v.ctypes.data_as(C.c_void_p),
                                    k.ctypes.data_as(C.c_void_p), bv.ctypes.data_as(C.c_void_p),
                                    sh.ctypes.data_as(C.c_void_p), tv.ctypes.data_as(C.c_void_p)))


class EulerParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("P0", C.c_double), ("T0", C.c_double), ("cp", C.c_double),
                ("cv", C.c_double), ("gravity", C.c_double * 3), ("buoyancy", C.c_int32),
                ("diffusion", C.c_int32), ("viscosity", C.c_double), ("Pr", C.c_double)]


BC_NEUMANN, BC_SYMMETRY, BC_DIRICHLET, BC_ROBIN = 0, 1, 2, 3


def euler_params(dt, P0=101325.0, T0=300.0, cp=1004.67, cv=715.5, gravity=(0.0, 0.0, -9.80606),
                 buoyancy=1, diffusion=0, viscosity=1.568e-5, Pr=0.9):
    """nsem_euler_params; viscosity and Pr default to apps/utils/properties.cpp:15-16."""
    p = EulerParams()
    p.dt, p.P0, p.T0, p.cp, p.cv = dt, P0, T0, cp, cv
    p.gravity[:] = list(gravity)
    p.buoyancy = buoyancy
    p.diffusion, p.viscosity, p.Pr = diffusion, viscosity, Pr
    return p


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DGEuler:
    """Device-resident dGSEM Euler problem (nsem_dg_* in the C-ABI)."""

    def __init__(self, mesh, params, bc_kind, bc_value):
        """bc_kind: (nb, 4) kinds of rho U T p; bc_value: (nb, 6) rho Ux Uy Uz T p."""
        self.mesh = mesh
        self.params = params
        k = np.ascontiguousarray(bc_kind, dtype=np.int32).reshape(-1)
        v = np.ascontiguousarray(bc_value, dtype=np.float64).reshape(-1)
        h = C.c_void_p()
        check(lib().nsem_dg_create(mesh.handle, C.byref(params), k.ctypes.data_as(C.c_void_p),
                                   v.ctypes.data_as(C.c_void_p), C.byref(h)))
        self._h = h
        self.nb = mesh.info.nbcs_field

    @classmethod
    def from_fields(cls, mesh, params, fields):
        """fields: dict name -> read_field() result for rho, U, T, p."""
        nb = len(mesh.boundaries)
        kinds = np.stack([fields[f][1] for f in ("rho", "U", "T", "p")], axis=1)
        vals = np.concatenate([fields["rho"][2], fields["U"][2], fields["T"][2], fields["p"][2]], axis=1)
        assert vals.shape == (nb, 6)
        return cls(mesh, params, kinds, vals)

    def set_state(self, rho, U, T, p):
        nb = self.nb
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (rho, U, T, p)]
        assert a[0].size == nb and a[1].size == 3 * nb and a[2].size == nb and a[3].size == nb
        check(lib().nsem_dg_set_state(self._h, *[_pd(x) for x in a]))
        self._keep = a

    def get_state(self):
        nb = self.nb
        rho, U, T, p = np.empty(nb), np.empty(3 * nb), np.empty(nb), np.empty(nb)
        check(lib().nsem_dg_get_state(self._h, _pd(rho), _pd(U), _pd(T), _pd(p)))
        return rho, U, T, p

    def affine(self):
        """True when the method-of-lines kernels use one metric per element."""
        v = C.c_int32()
        check(lib().nsem_dg_affine(self._h, C.byref(v)))
        return bool(v.value)

    def background(self):
        pr, rr = np.empty(self.nb), np.empty(self.nb)
        check(lib().nsem_dg_get_background(self._h, _pd(pr), _pd(rr)))
        return pr, rr

    def initial_density(self, T, p):
        T = np.ascontiguousarray(T, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        rho = np.empty(self.nb)
        check(lib().nsem_euler_initial_density(self._h, _pd(T), _pd(p), _pd(rho)))
        return rho

    def split_steps(self, n):
        check(lib().nsem_dg_split_steps(self._h, int(n)))

    def rk3_steps(self, n, precision=64):
        check(lib().nsem_dg_rk3_steps(self._h, int(n), int(precision)))

    def time_rk3(self, n, precision=64):
        s = C.c_double()
        check(lib().nsem_dg_time_rk3(self._h, int(n), int(precision), C.byref(s)))
        return s.value

    def time_split(self, n):
        s = C.c_double()
        check(lib().nsem_dg_time_split(self._h, int(n), C.byref(s)))
        return s.value

    def get_conserved(self):
        nb = self.nb
        rp, m, tp = np.empty(nb), np.empty(3 * nb), np.empty(nb)
        check(lib().nsem_dg_get_conserved(self._h, _pd(rp), _pd(m), _pd(tp)))
        return rp, m, tp

    # ---- derivatives of the FP64 method-of-lines step (conserved arrays of
    # shape (5, nb): rho', m_x, m_y, m_z, Theta')
    def residual_tlm(self, dq):
        dq = np.ascontiguousarray(dq, dtype=np.float64).reshape(5, -1)
        L, dL = np.empty_like(dq), np.empty_like(dq)
        check(lib().nsem_dg_residual_tlm(self._h, _pd(dq), _pd(L), _pd(dL)))
        return L, dL

    def residual_vjp(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(5, -1)
        t = np.empty_like(w)
        check(lib().nsem_dg_residual_vjp(self._h, _pd(w), _pd(t)))
        return t

    def rk3_tlm(self, n, dq):
        dq = np.array(dq, dtype=np.float64).reshape(5, -1)
        check(lib().nsem_dg_rk3_tlm(self._h, int(n), _pd(dq)))
        return dq

    def rk3_adjoint(self, n, lam, checkpoint=10):
        lam = np.array(lam, dtype=np.float64).reshape(5, -1)
        check(lib().nsem_dg_rk3_adjoint(self._h, int(n), int(checkpoint), _pd(lam)))
        return lam

    def time_rk3_adjoint(self, n, lam, checkpoint=10):
        """rk3_adjoint, also returning the device time of the sweeps."""
        lam = np.array(lam, dtype=np.float64).reshape(5, -1)
        s = C.c_double()
        check(lib().nsem_dg_time_rk3_adjoint(self._h, int(n), int(checkpoint), _pd(lam), C.byref(s)))
        return lam, s.value

    def set_conserved(self, rho_p, m, theta_p):
        """Set the method-of-lines state from conserved perturbations."""
        check(lib().nsem_dg_set_conserved(self._h, _pd(np.ascontiguousarray(rho_p, dtype=np.float64)),
                                          _pd(np.ascontiguousarray(m, dtype=np.float64)),
                                          _pd(np.ascontiguousarray(theta_p, dtype=np.float64))))

    def diagnostics(self):
        """(Courant max, min, avg, mass, energy, volume) of the current state."""
        out = np.empty(6)
        check(lib().nsem_dg_diagnostics(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def diagnostics_noalloc(self):
        """diagnostics() through the entry point that allocates nothing."""
        out = np.empty(6)
        check(lib().nsem_dg_diagnostics_noalloc(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def synchronize(self):
        check(lib().nsem_dg_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().nsem_dg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FvResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer_iterations", C.c_int32), ("breakdown", C.c_int32),
                ("reserved", C.c_int32), ("res0_ref", C.c_double), ("res_ref", C.c_double),
                ("res_true", C.c_double), ("seconds", C.c_double)]


class FVPoisson:
    """Device FV pressure Poisson (nsem_fv_* in the C-ABI)."""

    MODE_REFERENCE, MODE_TRUE_RESIDUAL, MODE_MIXED = 0, 1, 2

    def __init__(self, mesh, p_bc_kind, p_bc_value, mu=None):
        k = np.ascontiguousarray(p_bc_kind, dtype=np.int32)
        v = np.ascontiguousarray(p_bc_value, dtype=np.float64).reshape(-1)
        mu_p = None
        if mu is not None:
            self._mu = np.ascontiguousarray(mu, dtype=np.float64)
            mu_p = _pd(self._mu)
        h = C.c_void_p()
        check(lib().nsem_fv_create(mesh.handle, mu_p, k.ctypes.data_as(C.c_void_p),
                                   v.ctypes.data_as(C.c_void_p), C.byref(h)))
        self._h = h
        self.mesh = mesh
        n = C.c_uint64()
        st = C.c_int32()
        check(lib().nsem_fv_info(self._h, C.byref(n), C.byref(st)))
        self.n = n.value
        self.structured = bool(st.value)
        self.uniform = st.value == 2

    def set_rhs_divergence(self, U, bc_kind, bc_value):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        assert U.size == 3 * self.n
        k = np.ascontiguousarray(bc_kind, dtype=np.int32)
        v = np.ascontiguousarray(bc_value, dtype=np.float64).reshape(-1)
        check(lib().nsem_fv_set_rhs_divergence(self._h, _pd(U), k.ctypes.data_as(C.c_void_p), _pd(v),
                                               len(k)))

    def set_rhs(self, su):
        su = np.ascontiguousarray(su, dtype=np.float64)
        check(lib().nsem_fv_set_rhs(self._h, _pd(su)))

    def rhs(self):
        out = np.empty(self.n)
        check(lib().nsem_fv_get_rhs(self._h, _pd(out)))
        return out

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        check(lib().nsem_fv_apply(self._h, _pd(x), _pd(y)))
        return y

    def correct_velocity(self, p, U):
        """U - gradf(p, true) on the interior cells (nsem_fv_correct_velocity);
        U: (n, 3), returned as a new array."""
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)[:self.n]
        out = np.array(U, dtype=np.float64).reshape(-1)[:3 * self.n].copy()
        check(lib().nsem_fv_correct_velocity(self._h, _pd(p), _pd(out)))
        return out.reshape(-1, 3)

    def solve(self, x0=None, mode=0, tol=1e-10, max_it=100000, raise_on_error=True):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        res = FvResult()
        rc = lib().nsem_fv_solve(self._h, _pd(x), int(mode), C.c_double(tol), int(max_it), C.byref(res))
        if raise_on_error:
            check(rc)
        return x, res, rc

    def close(self):
        if getattr(self, "_h", None):
            lib().nsem_fv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
