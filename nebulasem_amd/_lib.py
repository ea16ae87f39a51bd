"""ctypes binding of libnebulasem_amd.so (the C-ABI in include/nebulasem_amd.h).

The library must have been built (nebulasem_amd/build.py, or
__graft_entry__.build()); nothing here falls back to Python or to the CPU.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnebulasem_amd.so")

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
