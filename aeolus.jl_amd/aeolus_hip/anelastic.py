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


def _buoyancy_frequency_squared(t_below, t_above, dz, g):
    """buoyancy_frequency_squared, logarithmic form (src/TurbulenceStability.jl:134-135)."""
    return g * (math.log(t_above) - math.log(t_below)) / dz


def _brunt_vaisala(theta, dz, g):
    """_brunt_vaisala (src/Anelastic.jl:326-338): N² = g ∂(log θ̄)/∂z,
    one-sided at the ends, centred over 2Δz inside."""
    Nz = len(theta)
    n2 = np.zeros(Nz)
    if Nz == 1:
        return n2
    n2[0] = _buoyancy_frequency_squared(theta[0], theta[1], dz, g)
    for k in range(1, Nz - 1):
        n2[k] = _buoyancy_frequency_squared(theta[k - 1], theta[k + 1], 2 * dz, g)
    n2[Nz - 1] = _buoyancy_frequency_squared(theta[Nz - 2], theta[Nz - 1], dz, g)
    return n2


class ReferenceState:
    """Cell and face profiles of a base state and their padded twins
    (_assemble_reference, src/Anelastic.jl:346-405): ρ̄, θ̄, p̄, Exner π̄ =
    (p̄/p₀)^(R/cₚ), T̄ = θ̄π̄, N², the pseudo-incompressible weight P = ρ̄θ̄;
    faces (log mean ρ̄, arithmetic θ̄, P_f = ρ̄_f θ̄_f) and padded twins."""

    def __init__(self, Nz, dz, H, theta, p, rho, R_d=R_DRY, cp=CP_DRY, p0=P_REF, g=G_STD):
        self.Nz, self.dz, self.H, self.g = Nz, dz, H, g
        self.R_d, self.cp, self.p0 = R_d, cp, p0
        self.rho, self.theta, self.p = (np.asarray(a, dtype=np.float64) for a in (rho, theta, p))
        if not (np.all(self.rho > 0) and np.all(self.theta > 0)):
            raise ValueError("reference density and potential temperature must be strictly positive")
        self.pi = np.array([(pk / p0) ** (R_d / cp) for pk in self.p])
        self.T = self.theta * self.pi
        self.N2 = _brunt_vaisala(self.theta, dz, g)
        self.P = self.rho * self.theta
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
        self.P_f = rf * tf
        self.Pc_pad = self.rho_c_pad * self.theta_c_pad
        self.Pf_pad = self.rho_f_pad * self.theta_f_pad

    def weights(self, weighting="anelastic"):
        """(Wc, Wf) interior profiles and their padded twins of a mass
        weighting: W = ρ̄ (anelastic) or ρ̄θ̄ (pseudo-incompressible);
        _weight_c/_weight_f (src/Anelastic.jl:1529-1532)."""
        if weighting == "anelastic":
            return self.rho, self.rho_f, self.rho_c_pad, self.rho_f_pad
        if weighting == "pseudo_incompressible":
            return self.P, self.P_f, self.Pc_pad, self.Pf_pad
        raise ValueError(f"unknown weighting {weighting!r} (anelastic, pseudo_incompressible)")

    def hydrostatic_residual(self):
        """(p̄[k] − p̄[k−1])/Δz + g ρ̄_f[k] at the interior faces
        (hydrostatic_residual, src/Anelastic.jl:627-635)."""
        return (self.p[1:] - self.p[:-1]) / self.dz + self.g * self.rho_f[1:self.Nz]

    def max_hydrostatic_residual(self):
        r = self.hydrostatic_residual()
        return float(np.abs(r).max()) if r.size else 0.0

    def scale_height(self):
        """reference_scale_height (src/Anelastic.jl:655-660)."""
        if self.Nz < 2:
            return math.inf
        return -((self.Nz - 1) * self.dz) / math.log(self.rho[-1] / self.rho[0])


def _density(p, theta, R, cp, p0):
    """_density (src/Anelastic.jl:312-313): ρ = (p₀/(Rθ)) (p/p₀)^{cᵥ/cₚ}."""
    return (p0 / (R * theta)) * (p / p0) ** ((cp - R) / cp)


def _integrate_hydrostatic(Nz, dz, theta, p_surf, R_d, cp, p0, g):
    """_integrate_hydrostatic (src/Anelastic.jl:409-445): level 1 by the
    half-cell balance, then the two-point balance with the log-mean face
    density, each by fixed-point iteration to 8 eps."""
    eps = np.finfo(np.float64).eps
    pb = np.empty(Nz)
    rb = np.empty(Nz)
    p = p_surf
    for _ in range(200):
        rho = _density(p, theta[0], R_d, cp, p0)
        pn = p_surf - g * (dz / 2) * rho
        d = abs(pn - p)
        p = pn
        if d <= 8 * eps * abs(p):
            break
    pb[0] = p
    rb[0] = _density(p, theta[0], R_d, cp, p0)
    for k in range(1, Nz):
        p = pb[k - 1] - g * dz * rb[k - 1]
        for _ in range(200):
            rho = _density(p, theta[k], R_d, cp, p0)
            pn = pb[k - 1] - g * dz * _logmean(rb[k - 1], rho)
            if not pn > 0:
                raise ValueError(f"hydrostatic integration produced non-positive pressure at k={k + 1}; "
                                 "domain is too deep for the prescribed θ̄ / p_surf")
            d = abs(pn - p)
            p = pn
            if d <= 8 * eps * abs(p):
                break
        pb[k] = p
        rb[k] = _density(p, theta[k], R_d, cp, p0)
    return pb, rb


def isentropic_reference(Nz, dz, H, theta_const=300.0, p_surf=P_REF, R_d=R_DRY, cp=CP_DRY, p0=P_REF, g=G_STD):
    """isentropic_reference (src/Anelastic.jl:493-504): θ̄ constant, p̄ and ρ̄
    from the discrete hydrostatic integration."""
    theta = np.full(Nz, float(theta_const))
    pb, rb = _integrate_hydrostatic(Nz, dz, theta, float(p_surf), R_d, cp, p0, g)
    return ReferenceState(Nz, dz, H, theta, pb, rb, R_d, cp, p0, g)


def constant_stability_reference(Nz, dz, H, N=0.01, theta_surf=300.0, z0=0.0, p_surf=P_REF, R_d=R_DRY, cp=CP_DRY,
                                 p0=P_REF, g=G_STD):
    """constant_stability_reference (src/Anelastic.jl:521-536): θ̄ = θ_s e^{N²ζ/g}."""
    n2 = N * N
    zeta = (z0 + (np.arange(Nz) + 0.5) * dz) - z0
    theta = np.array([theta_surf * math.exp(n2 * z / g) for z in zeta])
    pb, rb = _integrate_hydrostatic(Nz, dz, theta, float(p_surf), R_d, cp, p0, g)
    return ReferenceState(Nz, dz, H, theta, pb, rb, R_d, cp, p0, g)


def reference_from_profile(Nz, dz, H, theta_profile, p_surf=P_REF, z0=0.0, R_d=R_DRY, cp=CP_DRY, p0=P_REF, g=G_STD):
    """reference_from_profile (src/Anelastic.jl:551-568): θ̄ from a function
    of height (sampled at the cell centres) or an Nz-vector, integrated
    hydrostatically."""
    if callable(theta_profile):
        theta = np.array([float(theta_profile(z)) for z in z0 + (np.arange(Nz) + 0.5) * dz])
    else:
        theta = np.asarray(theta_profile, dtype=np.float64)
        if theta.shape != (Nz,):
            raise ValueError(f"θ_profile has length {theta.size}, expected Nz={Nz}")
    pb, rb = _integrate_hydrostatic(Nz, dz, theta, float(p_surf), R_d, cp, p0, g)
    return ReferenceState(Nz, dz, H, theta, pb, rb, R_d, cp, p0, g)


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

    weighting = "anelastic"

    def __init__(self, grid, ref, nu=0.0, flux_bot=0.0, flux_top=0.0, buoyancy=True,
                 bcs=("periodic", "periodic", "wall"), theta_bcs=("periodic", "periodic", "noflux"), precision=L.F64,
                 f_coriolis=0.0, sponge_alpha_max=0.0, sponge_z_start=None, sponge_target=None, C_s=0.0,
                 Pr_t=1 / 3):
        if grid.zf is not None:
            raise ValueError("the sound-proof core runs on regular grids")
        self.grid, self.ref, self.precision = grid, ref, precision
        self.bcs, self.theta_bcs = bcs, theta_bcs
        N = (grid.Nx, grid.Ny, grid.Nz)
        lat = tuple("periodic" if b == "periodic" else "neumann" for b in bcs[:2])
        wc, wf, wc_pad, wf_pad = ref.weights(self.weighting)
        self.solver = ColumnTridiagonalSolver.density_weighted(N, (grid.dx, grid.dy), grid.dz, wc, wf,
                                                               bc_x=lat[0], bc_y=lat[1])
        from .fields import BC_NAMES
        sp = L.MWParams()
        sp.nu, sp.g_accel, sp.flux_bot, sp.flux_top = nu, ref.g, flux_bot, flux_top
        sp.buoyancy = int(buoyancy)
        sp.bc_x, sp.bc_y, sp.bc_z = (BC_NAMES[b] for b in bcs)
        sp.tbc_x, sp.tbc_y, sp.tbc_z = (BC_NAMES[b] for b in theta_bcs)
        sp.precision = precision
        self._prof = [np.ascontiguousarray(a, dtype=np.float64) for a in (wc_pad, wf_pad, ref.theta_f_pad)]
        if any(len(a) != grid.Nz + 2 * grid.H for a in self._prof):
            raise ValueError("reference profiles must be padded with the grid's halo")
        h = ctypes.c_void_p()
        pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        L.check(L.init().aeo_mw_stepper_create(grid.handle, ctypes.byref(sp), *[pd(a) for a in self._prof],
                                               self.solver.handle, ctypes.byref(h)))
        self._h = h.value
        self.u, self.v, self.w, self.theta, self.p = (DeviceArray(grid.storage, precision) for _ in range(5))
        # device copies of the padded weights for the diagnostic divergence (one spare entry)
        self._wc = DeviceArray.from_numpy(np.append(wc_pad, wc_pad[-1]))
        self._wf = DeviceArray.from_numpy(np.append(wf_pad, wf_pad[-1]))
        self.t = 0.0
        self._parameters = dict(nu=float(nu), g_accel=float(ref.g), flux_bot=float(flux_bot),
                                flux_top=float(flux_top), buoyancy=bool(buoyancy), bcs=list(bcs),
                                theta_bcs=list(theta_bcs), f_coriolis=float(f_coriolis),
                                sponge_alpha_max=float(sponge_alpha_max),
                                sponge_z_start=None if sponge_z_start is None else float(sponge_z_start),
                                C_s=float(C_s), Pr_t=float(Pr_t))
        if C_s:
            L.check(L.init().aeo_mw_stepper_set_sgs(ctypes.c_void_p(self._h), float(C_s), float(Pr_t)))
        if f_coriolis:
            L.check(L.init().aeo_mw_stepper_set_coriolis(ctypes.c_void_p(self._h), float(f_coriolis)))
        z_top = grid.z0 + grid.Lz
        z_start = z_top if sponge_z_start is None else float(sponge_z_start)
        if sponge_alpha_max != 0 and z_start < z_top:
            from .simulation import damping_coordinates, sponge_levels
            zc, zf = damping_coordinates(grid)
            self._sponge = [zc, zf, *sponge_levels(sponge_target, grid.Nz)]
            pdn = lambda a: None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            L.check(L.init().aeo_mw_stepper_set_sponge(ctypes.c_void_p(self._h), float(sponge_alpha_max), z_start,
                                                       z_top, *[pdn(a) for a in self._sponge]))

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
        """∇·(Wu) over the interior, W the core's weight (mass_flux_divergence!,
        src/Anelastic.jl:1118-1123)."""
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


class PseudoIncompressibleSimulation(AnelasticSimulation):
    """The pseudo-incompressible (W = ρ̄θ̄, Durran) sound-proof core:
    MassWeightedSimulation{PseudoIncompressibleWeighting}
    (src/Anelastic.jl:1373, 1523-1532), the same step with the weights
    Pc_pad, Pf_pad."""

    weighting = "pseudo_incompressible"
