"""Device storage of fields and grids (src/Grids.jl, src/Fields.jl, src/StretchedGrids.jl).

A field is a DeviceArray of the reference's storage shape
(Nx+2H, Ny+2H, Nz+2H), column-major (x fastest); on the host side it maps to
a Fortran-ordered numpy array indexed a[i, j, k] like the Julia array.
"""
import ctypes

import numpy as np

from . import _lib as L

_DT = {L.F64: np.float64, L.F32: np.float32}


class DeviceArray:
    """Device memory from aeo_malloc, freed with the object."""

    def __init__(self, shape, precision=L.F64):
        self.shape = tuple(int(s) for s in shape)
        self.precision = precision
        self.dtype = np.dtype(_DT[precision])
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        lib = L.init()
        p = ctypes.c_void_p()
        L.check(lib.aeo_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value
        L.check(lib.aeo_memset(ctypes.c_void_p(self.ptr), 0, self.nbytes))

    @classmethod
    def from_numpy(cls, a, precision=None):
        a = np.asarray(a)
        if precision is None:
            precision = L.F32 if a.dtype == np.float32 else L.F64
        d = cls(a.shape, precision)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asfortranarray(np.asarray(a, dtype=self.dtype))
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        L.check(L.init().aeo_memcpy_h2d(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        L.check(L.init().aeo_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), self.nbytes))
        return out

    def zero(self):
        L.check(L.init().aeo_memset(ctypes.c_void_p(self.ptr), 0, self.nbytes))

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and L._lib is not None and L._initialised:
                L._lib.aeo_free(ctypes.c_void_p(self.ptr))
                self.ptr = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# vertical face generators (src/StretchedGrids.jl:117-278), host FP64
# ---------------------------------------------------------------------------
def faces_uniform(Nz, z0, Lz):
    """_faces_uniform (src/StretchedGrids.jl:133-136)."""
    dz = Lz / Nz
    return np.array([z0 + k * dz for k in range(Nz + 1)])


def faces_geometric(Nz, z0, Lz, r):
    """_faces_geometric (src/StretchedGrids.jl:151-166): dz[k+1] = r dz[k] filling Lz."""
    if abs(r - 1.0) <= 1e-12:
        return faces_uniform(Nz, z0, Lz)
    dz1 = Lz * (r - 1) / (r ** Nz - 1)
    zf = np.empty(Nz + 1)
    zf[0] = z0
    dz = dz1
    for k in range(Nz):
        zf[k + 1] = zf[k] + dz
        dz *= r
    zf[Nz] = z0 + Lz
    return zf


class Grid:
    """A RegularGrid (zf=None) or a StretchedGrid with interior faces zf (Nz+1)."""

    def __init__(self, size, extent, halo=3, zf=None, origin=(0.0, 0.0, 0.0)):
        self.Nx, self.Ny, self.Nz = (int(s) for s in size)
        self.H = int(halo)
        self.Lx, self.Ly, self.Lz = (float(e) for e in extent)
        self.x0, self.y0, self.z0 = (float(o) for o in origin)
        self.dx = self.Lx / self.Nx
        self.dy = self.Ly / self.Ny
        self.dz = self.Lz / self.Nz
        self.zf = None if zf is None else np.asarray(zf, dtype=np.float64)
        if self.zf is not None and len(self.zf) != self.Nz + 1:
            raise ValueError(f"custom face vector must have length Nz+1 = {self.Nz + 1}")
        self._h = None

    @property
    def storage(self):
        return (self.Nx + 2 * self.H, self.Ny + 2 * self.H, self.Nz + 2 * self.H)

    @property
    def interior(self):
        H = self.H
        return (slice(H, H + self.Nx), slice(H, H + self.Ny), slice(H, H + self.Nz))

    @property
    def handle(self):
        if self._h is None:
            lib = L.init()
            h = ctypes.c_void_p()
            zp = None
            if self.zf is not None:
                self._zf_c = np.ascontiguousarray(self.zf)
                zp = self._zf_c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            L.check(lib.aeo_grid_create(self.Nx, self.Ny, self.Nz, self.H, self.dx, self.dy, self.dz, zp,
                                        ctypes.byref(h)))
            self._h = h.value
        return ctypes.c_void_p(self._h)

    def metrics(self):
        """(idzc, idzf) of the storage levels, as the library built them."""
        n = self.Nz + 2 * self.H
        a = np.zeros(n)
        b = np.zeros(n + 1)
        L.check(L.init().aeo_grid_metrics(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a, b

    def interior_dz(self):
        """(dzc, dzf) over the interior: Nz cell thicknesses and Nz+1 centre spacings."""
        if self.zf is None:
            return np.full(self.Nz, self.dz), np.full(self.Nz + 1, self.dz)
        idzc, idzf = self.metrics()
        H = self.H
        return 1.0 / idzc[H:H + self.Nz], 1.0 / idzf[H:H + self.Nz + 1]

    def field(self, precision=L.F64):
        return DeviceArray(self.storage, precision)

    def __del__(self):
        try:
            if self._h is not None and L._lib is not None and L._initialised:
                L._lib.aeo_grid_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass


BC_NAMES = {"periodic": L.PERIODIC, "noflux": L.NOFLUX, "wall": L.WALL}


def fill_halos(grid, field, loc, bcs):
    """fill_halos!(field, bcs) (src/BoundaryConditions.jl:182-209); bcs = (bc_x, bc_y, bc_z)."""
    bx, by, bz = (BC_NAMES[b] if isinstance(b, str) else b for b in bcs)
    L.check(L.init().aeo_fill_halos(grid.handle, field.precision, L.dptr(field), loc, bx, by, bz))
    return field
