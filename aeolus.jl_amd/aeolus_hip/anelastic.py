"""The sound-proof (mass-weighted) core: reference states and the anelastic
step on the GPU.

ReferenceState profiles are built on the host, as the reference builds them
at construction (src/Anelastic.jl:211-388): cell profiles, face profiles (log
mean for ρ̄, arithmetic for θ̄, wall faces mirroring the adjacent cell), and
their halo-padded twins (_pad_center, _pad_face). The step itself,
_mass_weighted_step! (:1791-1876) with the anelastic weighting, runs in the
library's kernels (aeo_mw_step).
"""
import ctypes
import math

import numpy as np

from . import _lib as L
from .fields import DeviceArray, fill_halos
from .pressure import ColumnTridiagonalSolver

R_DRY = 287.0
CP_DRY = 1004.0
P_REF = 1.0e5
G_STD = 9.81


def _logmean(a, b):
    """_logmean (src/Anelastic.jl:254-267): (a − b)/log(a/b), series near a = b."""
    if a <= 0 or b <= 0:
        return (a + b) / 2
    s = a + b
    d = a - b
    if d == 0:
        return a
    x = d / s
    if abs(x) < 1e-3:
        x2 = x * x
        return (s / 2) * (1 - x2 / 3 - (4 * x2 * x2) / 45)
    return d / math.log(a / b)


def _ipow(x, n):
    r = x
    for _ in range(n - 1):
        r = r * x
    return r


def _pad_center(vals, H):
    Nz = len(vals)
    out = np.empty(Nz + 2 * H)
    out[H:H + Nz] = vals
    rb = vals[0] / vals[1] if Nz >= 2 else 1.0
    rt = vals[-1] / vals[-2] if Nz >= 2 else 1.0
    for h in range(1, H + 1):
        out[H - h] = vals[0] * _ipow(rb, h)
        out[H + Nz - 1 + h] = vals[-1] * _ipow(rt, h)
    return out


def _pad_face(vals, H, Nz):
    n = len(vals)
    out = np.empty(Nz + 2 * H)
    out[H:H + Nz + 1] = vals
    rb = vals[0] / vals[1] if n >= 2 else 1.0
    rt = vals[-1] / vals[-2] if n >= 2 else 1.0
    for h in range(1, H + 1):
        out[H - h] = vals[0] * _ipow(rb, h)
        idx = H + Nz + h
        if idx < len(out):
            out[idx] = vals[-1] * _ipow(rt, h)
    return out


class ReferenceState:
    """Cell and face profiles of a base state and their padded twins."""

    def __init__(self, Nz, dz, H, theta, p, rho, R_d=R_DRY, cp=CP_DRY, p0=P_REF, g=G_STD):
        self.Nz, self.dz, self.H, self.g = Nz, dz, H, g
        self.rho, self.theta, self.p = (np.asarray(a, dtype=np.float64) for a in (rho, theta, p))
        if not (np.all(self.rho > 0) and np.all(self.theta > 0)):
            raise ValueError("reference density and potential temperature must be strictly positive")
        rf = np.empty(Nz + 1)
        tf = np.empty(Nz + 1)
        rf[0], tf[0] = self.rho[0], self.theta[0]
        for k in range(1, Nz):
            rf[k] = _logmean(self.rho[k - 1], self.rho[k])
            tf[k] = (self.theta[k - 1] + self.theta[k]) / 2
        rf[Nz], tf[Nz] = self.rho[-1], self.theta[-1]
        self.rho_f, self.theta_f = rf, tf
        self.rho_c_pad = _pad_center(self.rho, H)
        self.theta_c_pad = _pad_center(self.theta, H)
        self.rho_f_pad = _pad_face(rf, H, Nz)
        self.theta_f_pad = _pad_face(tf, H, Nz)


def isothermal_reference(Nz, dz, H, z0=0.0, T_surf=288.0, p_surf=P_REF, R_d=R_DRY, cp=CP_DRY, p0=P_REF, g=G_STD):
    """isothermal_reference (src/Anelastic.jl:462-476): p̄ = p_s e^{−ζ/H_s},
    ρ̄ = p̄/(R T_s), θ̄ = T_s (p₀/p̄)^{R/c_p}, H_s = R T_s/g."""
    Hs = R_d * T_surf / g
    zeta = (z0 + (np.arange(Nz) + 0.5) * dz) - z0
    pb = p_surf * np.exp(-zeta / Hs)
    rho = pb / (R_d * T_surf)
    theta = T_surf * (p0 / pb) ** (R_d / cp)
    return ReferenceState(Nz, dz, H, theta, pb, rho, R_d, cp, p0, g)


class AnelasticSimulation:
    """The anelastic (W = ρ̄) sound-proof core on a regular grid: state on the
    device, step(dt) = _mass_weighted_step! (src/Anelastic.jl:1791-1876),
    projected with the density-weighted column solver (DirectWeightedPoissonSolver,
    src/Anelastic.jl:987-1050)."""

    def __init__(self, grid, ref, nu=0.0, flux_bot=0.0, flux_top=0.0, buoyancy=True,
                 bcs=("periodic", "periodic", "wall"), theta_bcs=("periodic", "periodic", "noflux"), precision=L.F64):
        if grid.zf is not None:
            raise ValueError("the sound-proof core runs on regular grids")
        self.grid, self.ref, self.precision = grid, ref, precision
        self.bcs, self.theta_bcs = bcs, theta_bcs
        N = (grid.Nx, grid.Ny, grid.Nz)
        lat = tuple("periodic" if b == "periodic" else "neumann" for b in bcs[:2])
        self.solver = ColumnTridiagonalSolver.density_weighted(N, (grid.dx, grid.dy), grid.dz, ref.rho, ref.rho_f,
                                                               bc_x=lat[0], bc_y=lat[1])
        from .fields import BC_NAMES
        sp = L.MWParams()
        sp.nu, sp.g_accel, sp.flux_bot, sp.flux_top = nu, ref.g, flux_bot, flux_top
        sp.buoyancy = int(buoyancy)
        sp.bc_x, sp.bc_y, sp.bc_z = (BC_NAMES[b] for b in bcs)
        sp.tbc_x, sp.tbc_y, sp.tbc_z = (BC_NAMES[b] for b in theta_bcs)
        sp.precision = precision
        self._prof = [np.ascontiguousarray(a, dtype=np.float64) for a in (ref.rho_c_pad, ref.rho_f_pad,
                                                                          ref.theta_f_pad)]
        if any(len(a) != grid.Nz + 2 * grid.H for a in self._prof):
            raise ValueError("reference profiles must be padded with the grid's halo")
        h = ctypes.c_void_p()
        pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        L.check(L.init().aeo_mw_stepper_create(grid.handle, ctypes.byref(sp), *[pd(a) for a in self._prof],
                                               self.solver.handle, ctypes.byref(h)))
        self._h = h.value
        self.u, self.v, self.w, self.theta, self.p = (DeviceArray(grid.storage, precision) for _ in range(5))
        # device copies of the padded weights for the diagnostic divergence (one spare entry)
        self._wc = DeviceArray.from_numpy(np.append(ref.rho_c_pad, ref.rho_c_pad[-1]))
        self._wf = DeviceArray.from_numpy(np.append(ref.rho_f_pad, ref.rho_f_pad[-1]))
        self.t = 0.0
        self._parameters = dict(nu=float(nu), g_accel=float(ref.g), flux_bot=float(flux_bot),
                                flux_top=float(flux_top), buoyancy=bool(buoyancy), bcs=list(bcs),
                                theta_bcs=list(theta_bcs))

    def parameters(self):
        """The run's scalar parameters, compared on restart."""
        return dict(self._parameters)

    def set(self, u, v, w, theta):
        for d, a in ((self.u, u), (self.v, v), (self.w, w), (self.theta, theta)):
            d.upload(a)

    def state(self):
        return dict(u=self.u.download(), v=self.v.download(), w=self.w.download(), theta=self.theta.download(),
                    p=self.p.download())

    def step(self, dt):
        L.check(L.init().aeo_mw_step(ctypes.c_void_p(self._h), L.dptr(self.u), L.dptr(self.v), L.dptr(self.w),
                                     L.dptr(self.theta), L.dptr(self.p), float(dt)))
        self.t += dt

    def step_rk3(self, dt):
        """step_rk3! of the sound-proof core (src/Anelastic.jl:1905-1945)."""
        L.check(L.init().aeo_mw_step_rk3(ctypes.c_void_p(self._h), L.dptr(self.u), L.dptr(self.v), L.dptr(self.w),
                                         L.dptr(self.theta), L.dptr(self.p), float(dt)))
        self.t += dt

    def mass_flux_divergence(self):
        """∇·(ρ̄u) over the interior (mass_flux_divergence!, src/Anelastic.jl:1118-1123)."""
        g = self.grid
        for a, loc in ((self.u, 1), (self.v, 2), (self.w, 3)):
            fill_halos(g, a, loc, self.bcs)
        out = DeviceArray(g.storage, self.precision)
        L.check(L.init().aeo_weighted_divergence(g.handle, self.precision, L.dptr(out), L.dptr(self.u),
                                                 L.dptr(self.v), L.dptr(self.w), L.dptr(self._wc),
                                                 L.dptr(self._wf), 0))
        return out.download()[g.interior]

    def __del__(self):
        try:
            if self._h and L._lib is not None and L._initialised:
                L._lib.aeo_mw_stepper_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass
