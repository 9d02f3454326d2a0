"""The Boussinesq dynamical core on the device: BoussinesqSimulation with
step_rk3! and step! (src/Simulations.jl:168-288, 474-487, 843-904).
"""
import ctypes
import math

import numpy as np

from . import _lib as L
from .advection import scheme_id
from .fields import BC_NAMES, Grid
from .pressure import FFTPoissonSolver, solver_for


def _bc(b):
    return BC_NAMES[b] if isinstance(b, str) else b


def damping_coordinates(grid):
    """The heights of the Nz interior levels at cell centres (u, v) and at z
    faces (w): damping_coordinates (src/FiniteVolume.jl:695, the RegularGrid
    znodes of src/Grids.jl:156-170; src/StretchedGrids.jl:372-378, 602)."""
    k = np.arange(1, grid.Nz + 1, dtype=np.float64)
    if grid.zf is None:
        return grid.z0 + (k - 0.5) * grid.dz, grid.z0 + (k - 1) * grid.dz
    zf = np.asarray(grid.zf, dtype=np.float64)
    return (zf[:-1] + zf[1:]) / 2, zf[:-1].copy()


def sponge_levels(target, Nz):
    """sponge_target(u, v, w) (src/Simulations.jl:319-352): a number or an
    Nz-level vector per component, absent components zero; returned as
    Nz-vectors (None for zero)."""
    target = {} if target is None else dict(target)
    for key in target:
        if key not in ("u", "v", "w"):
            raise ValueError(f"sponge_target: unknown component {key!r} (use u, v, w)")
    out = []
    for key in ("u", "v", "w"):
        t = target.get(key, 0)
        if np.isscalar(t):
            out.append(None if t == 0 else np.full(Nz, float(t)))
        else:
            t = np.ascontiguousarray(t, dtype=np.float64)
            if t.shape != (Nz,):
                raise ValueError(f"damping target: a level vector needs Nz = {Nz} values, got {t.size}")
            out.append(t)
    return out


class BoussinesqSimulation:
    """State u, v, w, theta, p on the device and one stepper.

    bcs / theta_bcs are (x, y, z) kinds: "periodic", "noflux" or "wall" (z
    only). The pressure solver follows the velocity BCs (_default_solver,
    src/Simulations.jl:1137-1144): FFT when every axis is periodic, else the
    column solver with the grid's z metrics.
    """

    def __init__(self, grid: Grid, scheme="centered", weno_p=2, nu=0.0, theta0=0.0, g_accel=9.81,
                 flux_bot=0.0, flux_top=0.0, bcs=("periodic",) * 3, theta_bcs=None, precision=L.F64,
                 solver=None, f_coriolis=0.0, sponge_alpha_max=0.0, sponge_z_start=None, sponge_target=None,
                 C_s=0.0, Pr_t=1 / 3):
        self.grid = grid
        self.precision = precision
        self.bcs = tuple(bcs)
        self.theta_bcs = tuple(theta_bcs) if theta_bcs is not None else self.bcs
        self.u = grid.field(precision)
        self.v = grid.field(precision)
        self.w = grid.field(precision)
        self.theta = grid.field(precision)
        self.p = grid.field(precision)
        if solver is None:
            if all(_bc(b) == L.PERIODIC for b in self.bcs):
                solver = FFTPoissonSolver(grid)
            else:
                solver = solver_for(grid)
        self.solver = solver
        bx, by, bz = (_bc(b) for b in self.bcs)
        tx, ty, tz = (_bc(b) for b in self.theta_bcs)
        sid = scheme_id(scheme)
        self.params = L.StepParams(sid, sid, int(weno_p), float(nu), float(theta0), float(g_accel),
                                   float(flux_bot), float(flux_top), bx, by, bz, tx, ty, tz, precision)
        h = ctypes.c_void_p()
        L.check(L.init().aeo_stepper_create(grid.handle, ctypes.byref(self.params), solver.handle, ctypes.byref(h)))
        self._h = h.value
        self.t = 0.0
        self._parameters = dict(scheme=scheme, weno_p=int(weno_p), nu=float(nu), theta0=float(theta0),
                                g_accel=float(g_accel), flux_bot=float(flux_bot), flux_top=float(flux_top),
                                f_coriolis=float(f_coriolis), sponge_alpha_max=float(sponge_alpha_max),
                                sponge_z_start=None if sponge_z_start is None else float(sponge_z_start),
                                sponge_target=None if sponge_target is None else
                                {k: np.asarray(v, dtype=np.float64).tolist() for k, v in dict(sponge_target).items()},
                                bcs=list(self.bcs), theta_bcs=list(self.theta_bcs), C_s=float(C_s),
                                Pr_t=float(Pr_t))
        if C_s:
            L.check(L.init().aeo_stepper_set_sgs(ctypes.c_void_p(self._h), float(C_s), float(Pr_t)))
        if f_coriolis:
            L.check(L.init().aeo_stepper_set_coriolis(ctypes.c_void_p(self._h), float(f_coriolis)))
        z_top = grid.z0 + grid.Lz
        z_start = z_top if sponge_z_start is None else float(sponge_z_start)
        if sponge_alpha_max != 0 and z_start < z_top:
            zc, zf = damping_coordinates(grid)
            tg = sponge_levels(sponge_target, grid.Nz)
            self._sponge = [zc, zf, *tg]
            pd = lambda a: None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            L.check(L.init().aeo_stepper_set_sponge(ctypes.c_void_p(self._h), float(sponge_alpha_max), z_start, z_top,
                                                    *[pd(a) for a in self._sponge]))

    def parameters(self):
        """The run's scalar parameters (_parameters, src/io/checkpoint.jl), compared on restart."""
        return dict(self._parameters)

    # ---- state ----
    def set(self, u=None, v=None, w=None, theta=None):
        for dev, a in ((self.u, u), (self.v, v), (self.w, w), (self.theta, theta)):
            if a is not None:
                dev.upload(a)

    def state(self):
        return {n: getattr(self, n).download() for n in ("u", "v", "w", "theta", "p")}

    # ---- steps ----
    def step_rk3(self, dt):
        L.check(L.init().aeo_step_rk3(ctypes.c_void_p(self._h), L.dptr(self.u), L.dptr(self.v), L.dptr(self.w),
                                      L.dptr(self.theta), L.dptr(self.p), float(dt)))
        self.t += dt

    def step(self, dt):
        L.check(L.init().aeo_step_euler(ctypes.c_void_p(self._h), L.dptr(self.u), L.dptr(self.v),
                                        L.dptr(self.w), L.dptr(self.theta), L.dptr(self.p), float(dt)))
        self.t += dt

    # ---- diagnostics ----
    def fill_velocity_halos(self):
        from .fields import fill_halos
        fill_halos(self.grid, self.u, L.FACE_X, self.bcs)
        fill_halos(self.grid, self.v, L.FACE_Y, self.bcs)
        fill_halos(self.grid, self.w, L.FACE_Z, self.bcs)

    def max_divergence(self):
        """max_divergence(sim) (src/Simulations.jl:1045-1050)."""
        self.fill_velocity_halos()
        out = ctypes.c_double()
        L.check(L.init().aeo_max_divergence(self.grid.handle, self.precision, L.dptr(self.u), L.dptr(self.v),
                                            L.dptr(self.w), ctypes.byref(out)))
        return out.value

    def kinetic_energy(self):
        """kinetic_energy(sim) (src/Simulations.jl:992-1015): ½ ∫(u² + v² + w²) dV,
        u, v, w as stored on their faces."""
        out = ctypes.c_double()
        L.check(L.init().aeo_kinetic_energy(self.grid.handle, self.precision, L.dptr(self.u), L.dptr(self.v),
                                            L.dptr(self.w), ctypes.byref(out)))
        return out.value

    def advective_dt(self, C=0.5):
        """dt = C / max Σ|uᵢ|/Δxᵢ (CFLConstraints C_advective, src/TimeIntegration.jl:127-182)."""
        self.fill_velocity_halos()
        out = ctypes.c_double()
        L.check(L.init().aeo_max_advective_rate(self.grid.handle, self.precision, L.dptr(self.u), L.dptr(self.v),
                                                L.dptr(self.w), ctypes.byref(out)))
        return C / out.value if out.value > 0 else np.inf

    def __del__(self):
        try:
            if self._h and L._lib is not None and L._initialised:
                L._lib.aeo_stepper_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass


class CFLConstraints:
    """CFLConstraints (src/TimeIntegration.jl:127-163): safety factors of the
    advective, diffusive, buoyancy and Coriolis limits, the dt clamp and the
    growth bound between consecutive steps."""

    def __init__(self, C_advective=0.5, C_diffusive=0.25, C_buoyancy=0.5, C_coriolis=0.5, dt_min=1e-8,
                 dt_max=math.inf, growth_max=1.1):
        if not C_advective > 0:
            raise ValueError("C_advective must be > 0")
        if not C_diffusive > 0:
            raise ValueError("C_diffusive must be > 0")
        if not C_buoyancy > 0:
            raise ValueError("C_buoyancy must be > 0")
        if not C_coriolis > 0:
            raise ValueError("C_coriolis must be > 0")
        if not 0 < dt_min <= dt_max:
            raise ValueError("need 0 < dt_min ≤ dt_max")
        if not growth_max >= 1:
            raise ValueError("growth_max must be ≥ 1")
        self.C_advective, self.C_diffusive, self.C_buoyancy = float(C_advective), float(C_diffusive), float(C_buoyancy)
        self.C_coriolis, self.dt_min, self.dt_max = float(C_coriolis), float(dt_min), float(dt_max)
        self.growth_max = float(growth_max)


def cfl_diagnostics(sim, c=None, limits=()):
    """cfl_diagnostics(sim, constraints; limits) (src/TimeIntegration.jl:291-400)
    with the simulation's own closure and sgs_stress = :laplacian: the limit of
    each process, the binding one and the maxima it came from. `limits` are
    (dt, label) pairs of further components."""
    c = c or CFLConstraints()
    rate, nut, N = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    L.check(L.init().aeo_stepper_cfl(ctypes.c_void_p(sim._h), L.dptr(sim.u), L.dptr(sim.v), L.dptr(sim.w),
                                     L.dptr(sim.theta), ctypes.byref(rate), ctypes.byref(nut), ctypes.byref(N)))
    return cfl_limits(rate.value, nut.value, N.value, sim.parameters(), sim.grid, c, limits)


def cfl_limits(rate, nut_max, N, p, g, c, limits=()):
    """The limits of cfl_diagnostics from the state maxima (the advective rate,
    max νt, N), the run parameters p (nu, C_s, Pr_t, f_coriolis) and the grid."""
    inf = math.inf
    dt_adv = math.nan if math.isnan(rate) else (c.C_advective / rate if rate > 0 else inf)
    kappa_max = nut_max / p["Pr_t"] if p["C_s"] else 0.0
    dz_min = g.dz if g.zf is None else float(np.min(np.diff(np.asarray(g.zf, dtype=np.float64))))
    d2harm = 1.0 / (1 / g.dx ** 2 + 1 / g.dy ** 2 + 1 / dz_min ** 2)
    nut_cfl = nut_max
    nu_total = p["nu"] + nut_cfl
    dt_diff = c.C_diffusive * d2harm / nu_total if nu_total > 0 else inf
    k_sgs = max(nut_cfl, kappa_max)
    dt_sgs = c.C_diffusive * d2harm / k_sgs if k_sgs > 0 else inf
    dt_buoy = c.C_buoyancy / N if N > 0 else inf
    f = abs(p["f_coriolis"])
    dt_cor = c.C_coriolis / f if f > 0 else inf
    dt_lim, limiter = dt_adv, "advective"
    for d, name in ((dt_diff, "diffusive"), (dt_buoy, "buoyancy"), (dt_cor, "coriolis")):
        if d < dt_lim:
            dt_lim, limiter = d, name
    dt_limits, lim_label = inf, "none"
    for lim in limits:
        if lim is None:
            continue
        d_l, label = float(lim[0]), lim[-1]
        if d_l < dt_limits:
            dt_limits, lim_label = d_l, label
        if d_l < dt_lim:
            dt_lim, limiter = d_l, label
    if dt_sgs < dt_lim:
        dt_lim, limiter = dt_sgs, "sgs"
    return dict(dt_adv=dt_adv, dt_diff=dt_diff, dt_buoy=dt_buoy, dt_cor=dt_cor, dt_sgs=dt_sgs, dt_limits=dt_limits,
                limits_limiter=lim_label, dt_limiting=dt_lim, limiter=limiter, advective_rate=rate, N=N,
                nut_max=nut_max, kappa_max=kappa_max, nut_cfl=nut_cfl, nu_total=nu_total)


def adaptive_dt(sim, c=None, dt_prev=None, diagnostics=None, limits=()):
    """adaptive_dt (src/TimeIntegration.jl:436-466): the binding limit, then the
    growth bound against dt_prev, then the clamp to [dt_min, dt_max]; a NaN
    limit collapses to dt_min. On a sound-proof core, that core's own method
    (src/Anelastic.jl:2240-2300)."""
    c = c or CFLConstraints()
    if diagnostics is None and getattr(sim, "weighting", None) is not None:
        from .timestepping import mw_adaptive_dt  # the sound-proof cores' method
        return mw_adaptive_dt(sim, c, dt_prev, limits)
    d = diagnostics or cfl_diagnostics(sim, c, limits)
    dt = d["dt_limiting"]
    if math.isnan(dt):
        return c.dt_min
    if dt_prev is not None and dt_prev > 0 and c.growth_max * dt_prev < dt:
        dt = c.growth_max * dt_prev
    return min(max(dt, c.dt_min), c.dt_max)
