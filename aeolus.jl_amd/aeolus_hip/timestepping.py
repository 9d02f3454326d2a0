"""The adaptive step and its failure handling (src/TimeIntegration.jl:470-940;
the sound-proof cores' methods, src/Anelastic.jl:2214-2370).

One call of step_adaptive picks dt (adaptive_dt), saves the prognostic state
into a snapshot kept on the simulation, takes the step, and checks it with
step_guard (one fused sweep on the GPU: finiteness, the constraint residual,
the speed). A step that throws or fails a check is rolled back from the
snapshot and retried at half the step, never below dt_min, up to
max_retries times; then the pre-step state is restored and the call raises.
Fast processes (callbacks) run afterwards, substepped by a
SubsteppingSchedule.

The snapshot holds u, v, w, p, θ and the clock. The eddy viscosity is not
part of it: every step and every CFL evaluation recomputes νt from the state
before it reads it.
"""
import ctypes
import logging
import math

from . import _lib as L
from .fields import DeviceArray

log = logging.getLogger("aeolus_hip.timestepping")


class NumericalFailure(RuntimeError):
    """NumericalFailure(msg) (src/TimeIntegration.jl:610-624): a step, a solver or
    a physics process that failed numerically (diverged, did not converge, left
    its physical range), as opposed to a programming error."""


def _is_mass_weighted(sim):
    return getattr(sim, "weighting", None) is not None


def step_guard(sim):
    """step_guard(sim) -> dict(finite, divergence, speed) (src/TimeIntegration.jl:
    673-710; src/Anelastic.jl:2333-2346 for the sound-proof cores, whose
    divergence is the constraint residual max |(1/W)∇·(W u)|). The velocity
    halos are filled first; divergence and speed propagate NaN."""
    fin, div, spd = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    fn = L.init().aeo_mw_stepper_guard if _is_mass_weighted(sim) else L.init().aeo_stepper_guard
    L.check(fn(ctypes.c_void_p(sim._h), L.dptr(sim.u), L.dptr(sim.v), L.dptr(sim.w), L.dptr(sim.theta),
               L.dptr(sim.p), ctypes.byref(fin), ctypes.byref(div), ctypes.byref(spd)))
    return dict(finite=bool(fin.value), divergence=div.value, speed=spd.value)


def _all_finite(sim):
    return step_guard(sim)["finite"]


def is_numerical_failure(err, sim):
    """is_numerical_failure(err, sim) (src/TimeIntegration.jl:626-654): a
    NumericalFailure, a domain or overflow error, or any exception raised while
    the prognostic state is no longer finite; never a KeyboardInterrupt."""
    if isinstance(err, KeyboardInterrupt):
        return False
    if isinstance(err, (NumericalFailure, ArithmeticError)):
        return True
    try:
        return not _all_finite(sim)
    except Exception:
        return False


class SubsteppingSchedule:
    """SubsteppingSchedule(default=1, **counts) (src/TimeIntegration.jl:476-492):
    how many substeps each named fast process takes per step."""

    def __init__(self, default=1, **counts):
        if int(default) < 1:
            raise ValueError("default substep count must be >= 1")
        for k, v in counts.items():
            if int(v) < 1:
                raise ValueError(f"substep count for {k} must be >= 1, got {v}")
        self.default = int(default)
        self.counts = {k: int(v) for k, v in counts.items()}


def substeps_for(schedule, name):
    """substeps_for(schedule, name) (src/TimeIntegration.jl:501): the schedule's
    count for `name`, its default when `name` is not registered; None is an
    empty schedule."""
    if schedule is None:
        return 1
    return schedule.counts.get(name, schedule.default)


class StateSnapshot:
    """StateSnapshot (src/TimeIntegration.jl:527-545): device copies of u, v, w,
    p, θ (with halos) and the clock, allocated once per simulation."""

    def __init__(self, sim):
        self.fields = {n: DeviceArray(getattr(sim, n).shape, sim.precision) for n in ("u", "v", "w", "p", "theta")}
        self.t = sim.t
        self.valid = False


def _copy(dst, src):
    L.check(L.init().aeo_memcpy_d2d(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), src.nbytes))


def save_state(snap, sim):
    """save_state!(snap, sim) (src/TimeIntegration.jl:553-561): device copies."""
    for n, a in snap.fields.items():
        _copy(a, getattr(sim, n))
    snap.t = sim.t
    snap.valid = True
    return snap


def restore_state(sim, snap):
    """restore_state!(sim, snap) (src/TimeIntegration.jl:563-580): the fields with
    their halos and the clock; synchronised before returning."""
    if not snap.valid:
        raise ValueError("StateSnapshot has never been saved")
    for n, a in snap.fields.items():
        _copy(getattr(sim, n), a)
    L.check(L.init().aeo_sync())
    sim.t = snap.t
    return sim


def state_snapshot(sim):
    """state_snapshot(sim) (src/TimeIntegration.jl:582-597): the simulation's
    snapshot buffer, allocated on first use and kept."""
    snap = getattr(sim, "_snapshot", None)
    if snap is None:
        snap = sim._snapshot = StateSnapshot(sim)
    return snap


def _fast_process_pairs(fp):
    """_fast_process_pairs (src/TimeIntegration.jl:727-732)."""
    if fp is None:
        return ()
    if callable(fp):
        return (("default", fp),)
    if isinstance(fp, dict):
        return tuple(fp.items())
    return tuple(fp)


def _guard_reason(sim, div_tol, u_cap):
    """_guard_reason (src/TimeIntegration.jl:713-721): non-finite state, then the
    divergence, then the speed cap."""
    g = step_guard(sim)
    if not g["finite"]:
        return "nonfinite"
    if not math.isfinite(g["divergence"]) or g["divergence"] > div_tol:
        return "divergence"
    if g["speed"] > u_cap:
        return "velocity"
    return "ok"


def step_adaptive(sim, constraints=None, dt=None, dt_prev=None, fast_processes=None, schedule=None, max_retries=3,
                  stepper=None, snapshot=None, divergence_tol=1e-6, max_velocity=math.inf, limits=()):
    """step_adaptive!(sim, constraints; ...) -> (sim, dt_used, n_retries)
    (src/TimeIntegration.jl:734-920).

    dt from adaptive_dt unless given; the state saved; the step taken with
    `stepper(sim, dt)` (default: sim.step, the reference's step!); a step that
    raises or fails step_guard (non-finite, divergence above divergence_tol,
    max speed above max_velocity) is rolled back and retried at
    max(dt/2, dt_min) up to max_retries times. When every retry fails, or dt
    cannot shrink further, the pre-step state is restored and a RuntimeError
    is raised (the reference's error(...)). Then each fast process runs
    substeps_for(schedule, name) times with dt/n."""
    from .simulation import CFLConstraints, adaptive_dt
    c = constraints or CFLConstraints()
    h = float(dt) if dt is not None else adaptive_dt(sim, c, dt_prev=dt_prev, limits=limits)
    take_step = stepper if stepper is not None else (lambda s, d: s.step(d))
    snap = snapshot if snapshot is not None else state_snapshot(sim)
    save_state(snap, sim)
    n_retries = 0
    while True:
        reason = "ok"
        try:
            take_step(sim, h)
        except (KeyboardInterrupt, SystemExit):
            raise
        except Exception as err:  # a throw the driver recovers from by a retry
            reason = "exception"
            log.debug("step_adaptive: stepper threw", extra=dict(subsystem="timestepping", dt=h, retry=n_retries,
                                                                  exception=repr(err)))
        if reason == "ok":
            reason = _guard_reason(sim, divergence_tol, max_velocity)
        if reason == "ok":
            break
        if n_retries >= max_retries:
            restore_state(sim, snap)
            raise RuntimeError(f"step_adaptive: step failed ({reason}) after {max_retries} retries; "
                                   f"last dt = {h}. State has been restored to the beginning of the step.")
        restore_state(sim, snap)
        h_new = max(h / 2, c.dt_min)
        if h_new == h:
            raise RuntimeError(f"step_adaptive: step failed ({reason}) at dt_min = {c.dt_min}; cannot reduce "
                                   f"dt further. State has been restored to the beginning of the step.")
        log.debug("step_adaptive: retry", extra=dict(subsystem="timestepping", reason=reason, dt_from=h, dt_to=h_new,
                                                     retry=n_retries + 1, max_retries=max_retries))
        h = h_new
        n_retries += 1
    procs = _fast_process_pairs(fast_processes)
    if procs:
        for name, cb in procs:
            n = substeps_for(schedule, name)
            for _ in range(n):
                cb(sim, h / n)
        if not _all_finite(sim):
            log.warning("step_adaptive: state became non-finite inside a fast process",
                        extra=dict(subsystem="timestepping", dt=h, processes=tuple(p[0] for p in procs)))
    return sim, h, n_retries


def mw_cfl_maxima(sim):
    """The state maxima of the sound-proof adaptive_dt: (rate, max νt, N)."""
    rate, nut, N = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    L.check(L.init().aeo_mw_stepper_cfl(ctypes.c_void_p(sim._h), L.dptr(sim.u), L.dptr(sim.v), L.dptr(sim.w),
                                        L.dptr(sim.theta), ctypes.byref(rate), ctypes.byref(nut), ctypes.byref(N)))
    return rate.value, nut.value, N.value


def mw_dt_pick(rate, nut_max, N, p, g, c, dt_prev=None, limits=()):
    """adaptive_dt(sim::MassWeightedSimulation, c) (src/Anelastic.jl:2240-2300)
    from the state maxima: the advective limit, the diffusive limit with
    ν + max(νt, κt), the buoyancy and Coriolis limits and the components'
    limits in that order; then the growth bound and the clamp. A NaN rate
    returns dt_min. Returns (dt, limiter)."""
    if math.isnan(rate):
        return c.dt_min, "nan"
    dt = c.C_advective / rate if rate > 0 else math.inf
    limiter = "advective"
    kmax = max(nut_max, nut_max / p["Pr_t"]) if p["C_s"] else 0.0
    nutot = p["nu"] + kmax
    d2harm = 1.0 / (1 / g.dx ** 2 + 1 / g.dy ** 2 + 1 / g.dz ** 2)
    if nutot > 0 and c.C_diffusive * d2harm / nutot < dt:
        dt, limiter = c.C_diffusive * d2harm / nutot, "diffusive"
    if N > 0 and c.C_buoyancy / N < dt:
        dt, limiter = c.C_buoyancy / N, "buoyancy"
    f = p["f_coriolis"]
    if f != 0 and c.C_coriolis / abs(f) < dt:
        dt, limiter = c.C_coriolis / abs(f), "coriolis"
    for lim in limits:
        if lim is None:
            continue
        d_l = float(lim[0])
        if d_l < dt:
            dt, limiter = d_l, lim[-1]
    if math.isnan(dt):
        return c.dt_min, "nan"
    if dt_prev is not None and dt_prev > 0 and c.growth_max * dt_prev < dt:
        dt, limiter = c.growth_max * dt_prev, "growth"
    if dt < c.dt_min:
        limiter = "dt_min"
    if dt > c.dt_max:
        limiter = "dt_max"
    return min(max(dt, c.dt_min), c.dt_max), limiter


def mw_adaptive_dt(sim, c, dt_prev=None, limits=()):
    rate, nut, N = mw_cfl_maxima(sim)
    return mw_dt_pick(rate, nut, N, sim.parameters(), sim.grid, c, dt_prev, limits)[0]


__all__ = ["NumericalFailure", "is_numerical_failure", "SubsteppingSchedule", "substeps_for", "StateSnapshot",
           "save_state", "restore_state", "state_snapshot", "step_guard", "step_adaptive", "mw_dt_pick"]
