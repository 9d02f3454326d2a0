"""Pressure solvers: the AbstractPressureSolver surface (src/PressureSolvers.jl:19-80).

    solve(p, source, solver)          solve!(p, source, solver)
    source_buffer(solver)             the FP64 interior-sized device buffer
    solve_source_buffer(phi, solver)  solve_source_buffer!(p, solver)
    solver_stats(solver)              (iterations, residual)
"""
import ctypes
import math

import numpy as np

from . import _lib as L
from .fields import DeviceArray


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class PressureSolver:
    kind = "abstract"

    def __init__(self, handle, size):
        self._h = handle
        self.Nx, self.Ny, self.Nz = size

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def __del__(self):
        try:
            if self._h and L._lib is not None and L._initialised:
                L._lib.aeo_poisson_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass


class FFTPoissonSolver(PressureSolver):
    """FFTPoissonSolver(mesh) (src/PressureSolvers.jl:103-174): triply periodic."""

    kind = "fft"
    axis_bcs = ("periodic", "periodic", "periodic")

    def __init__(self, grid):
        h = ctypes.c_void_p()
        L.check(L.init().aeo_poisson_fft_create(grid.Nx, grid.Ny, grid.Nz, grid.dx, grid.dy, grid.dz,
                                                ctypes.byref(h)))
        super().__init__(h.value, (grid.Nx, grid.Ny, grid.Nz))


class ColumnTridiagonalSolver(PressureSolver):
    """ColumnTridiagonalSolver (src/PressureSolvers.jl:544-805): periodic x-y
    transform + FP64 Thomas in z + mean mode, with the column coefficients
    lower, upper, hweight, measure (Nz) and inv_coupling (Nz+1), gauge (Nz)."""

    kind = "column"
    axis_bcs = ("periodic", "periodic", "neumann")

    def __init__(self, size, dxy, lower, upper, hweight, measure, inv_coupling, gauge, bc_x="periodic",
                 bc_y="periodic"):
        Nx, Ny, Nz = size
        vecs = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in
                (lower, upper, hweight, measure, inv_coupling, gauge)]
        for v, n in zip(vecs, (Nz, Nz, Nz, Nz, Nz + 1, Nz)):
            if len(v) != n:
                raise ValueError("column coefficients must have length Nz (inv_coupling: Nz + 1)")
        bx = L.PERIODIC if bc_x == "periodic" else L.NOFLUX
        by = L.PERIODIC if bc_y == "periodic" else L.NOFLUX
        h = ctypes.c_void_p()
        L.check(L.init().aeo_poisson_column_create(Nx, Ny, Nz, float(dxy[0]), float(dxy[1]), bx, by,
                                                   *[_pd(v) for v in vecs], ctypes.byref(h)))
        self._vecs = vecs
        super().__init__(h.value, (Nx, Ny, Nz))

    @classmethod
    def stretched(cls, size, dxy, dzc, dzf, **kw):
        """The conservative stretched-z Laplacian (src/PressureSolvers.jl:685-700):
        a_k = 1/(dzc[k] dzf[k]), c_k = 1/(dzc[k] dzf[k+1]), h = 1, measure = gauge = dzc."""
        Nz = size[2]
        zc = np.asarray(dzc, dtype=np.float64)
        zf = np.asarray(dzf, dtype=np.float64)
        if len(zc) != Nz or len(zf) != Nz + 1:
            raise ValueError("dzc must have length Nz and dzf length Nz + 1")
        lower = np.zeros(Nz)
        upper = np.zeros(Nz)
        for k in range(Nz):
            if k > 0:
                lower[k] = 1 / (zc[k] * zf[k])
            if k < Nz - 1:
                upper[k] = 1 / (zc[k] * zf[k + 1])
        return cls(size, dxy, lower, upper, np.ones(Nz), zc, zf, zc, **kw)

    @classmethod
    def density_weighted(cls, size, dxy, dz, w_c, w_f, **kw):
        """DirectWeightedPoissonSolver (src/Anelastic.jl:987-1040): ∇·(w(z)∇φ)."""
        Nz = size[2]
        wc = np.asarray(w_c, dtype=np.float64)
        wf = np.asarray(w_f, dtype=np.float64)
        if len(wc) != Nz or len(wf) != Nz + 1:
            raise ValueError("w_c must have length Nz and w_f length Nz+1")
        if not (np.all(wc > 0) and np.all(wf > 0)):
            raise ValueError("Poisson weights must be strictly positive")
        idz2 = 1 / dz ** 2
        lower = np.array([wf[k] * idz2 if k > 0 else 0.0 for k in range(Nz)])
        upper = np.array([wf[k + 1] * idz2 if k < Nz - 1 else 0.0 for k in range(Nz)])
        return cls(size, dxy, lower, upper, wc, np.full(Nz, dz ** 2), 1.0 / wf, np.ones(Nz), **kw)


def solver_for(grid):
    """The default solver of a grid (_default_solver, src/Simulations.jl:1137-1144, and
    StretchedPoissonSolver, src/stretched/poisson.jl:78-103) for periodic x-y."""
    dzc, dzf = grid.interior_dz()
    return ColumnTridiagonalSolver.stretched((grid.Nx, grid.Ny, grid.Nz), (grid.dx, grid.dy), dzc, dzf)


def source_buffer(solver):
    p = ctypes.c_void_p()
    L.check(L.init().aeo_poisson_source_buffer(solver.handle, ctypes.byref(p)))
    return p.value


def solve_source_buffer(phi, solver):
    """φ (FP64 interior-sized DeviceArray) from the source already in source_buffer(solver)."""
    L.check(L.init().aeo_poisson_solve_source_buffer(solver.handle, L.dptr(phi)))
    return phi


def solve(p, source, solver, H):
    """solve!(p, source, solver): storage-layout DeviceArrays of halo H."""
    L.check(L.init().aeo_poisson_solve(solver.handle, p.precision, L.dptr(p), L.dptr(source), H))
    return p


def solver_stats(solver):
    it = ctypes.c_int()
    r = ctypes.c_double()
    st = ctypes.c_int()
    L.check(L.init().aeo_poisson_stats(solver.handle, ctypes.byref(it), ctypes.byref(r), ctypes.byref(st)))
    return it.value, r.value


def solve_interior(solver, src):
    """Convenience: numpy (Nx, Ny, Nz) source -> numpy φ through the source buffer."""
    n = solver.Nx * solver.Ny * solver.Nz
    a = np.asfortranarray(np.asarray(src, dtype=np.float64))
    L.check(L.init().aeo_memcpy_h2d(ctypes.c_void_p(source_buffer(solver)), a.ctypes.data_as(ctypes.c_void_p),
                                    8 * n))
    phi = DeviceArray((solver.Nx, solver.Ny, solver.Nz), L.F64)
    solve_source_buffer(phi, solver)
    return phi.download()
