"""The control flow of step_adaptive (src/TimeIntegration.jl:808-920) on the
CPU: the device calls it makes (the snapshot copies, the fused guard) are
replaced by numpy stand-ins on a host-side state, so the retry/halving/
restore/substepping logic is checked without a GPU. The cases follow the
reference's test/timeint.jl:55-90 (fast-process substeps, a stiff growth
rejected twice, an exhausted retry budget that restores the state)."""
import math

import numpy as np
import pytest

import aeolus_hip.timestepping as ts
from aeolus_hip.simulation import CFLConstraints


class HostSim:
    """u uniform 1 on an 8^3 interior; step(dt) multiplies u by (1 + dt G)
    (the stiff explicit growth of the reference's blowup! body force)."""

    def __init__(self, G=0.0, poison=False):
        self.u = np.ones((8, 8, 8))
        self.v = np.zeros((8, 8, 8))
        self.w = np.zeros((8, 8, 8))
        self.p = np.zeros((8, 8, 8))
        self.theta = np.zeros((8, 8, 8))
        self.t, self.G, self.poison, self.calls = 0.0, G, poison, []

    def step(self, dt):
        self.calls.append(dt)
        self.u *= 1 + dt * self.G
        if self.poison:
            self.u[2, 2, 2] = np.nan
        self.t += dt


class HostSnap:
    def __init__(self):
        self.valid = False


@pytest.fixture
def host(monkeypatch):
    def save(snap, sim):
        snap.a = {n: getattr(sim, n).copy() for n in ("u", "v", "w", "p", "theta")}
        snap.t, snap.valid = sim.t, True
        return snap

    def restore(sim, snap):
        if not snap.valid:
            raise ValueError("StateSnapshot has never been saved")
        for n, a in snap.a.items():
            getattr(sim, n)[...] = a
        sim.t = snap.t
        return sim

    def guard(sim):
        fin = all(np.isfinite(getattr(sim, n)).all() for n in ("u", "v", "w", "p", "theta"))
        spd = max(np.abs(sim.u).max(), np.abs(sim.v).max(), np.abs(sim.w).max())
        return dict(finite=fin, divergence=0.0, speed=spd)

    monkeypatch.setattr(ts, "save_state", save)
    monkeypatch.setattr(ts, "restore_state", restore)
    monkeypatch.setattr(ts, "step_guard", guard)
    monkeypatch.setattr(ts, "state_snapshot", lambda sim: HostSnap())
    return ts


def test_fast_processes_are_substepped(host):
    sim = HostSim()
    got = []
    _, dt, nr = host.step_adaptive(sim, CFLConstraints(dt_max=0.05), dt=0.05,
                                   fast_processes=dict(microphysics=lambda s, h: got.append(h)),
                                   schedule=host.SubsteppingSchedule(microphysics=4))
    assert nr == 0 and dt == 0.05 and len(got) == 4 and all(h == pytest.approx(dt / 4) for h in got)
    # a bare callback is the :default process, one substep unless the schedule says otherwise
    got.clear()
    host.step_adaptive(sim, dt=0.01, fast_processes=lambda s, h: got.append(h),
                       schedule=host.SubsteppingSchedule(default=3))
    assert got == [pytest.approx(0.01 / 3)] * 3


def test_stiff_growth_rejected_twice_then_accepted(host):
    sim = HostSim(G=1.0e4)
    c = CFLConstraints(dt_min=1e-8, dt_max=1.0)
    # dt = 1e-3: |u| = 11 (reject); 5e-4: 6 (reject); 2.5e-4: 3.5 (accept)
    _, dt, nr = host.step_adaptive(sim, c, dt=1e-3, max_velocity=4.0, max_retries=3)
    assert nr == 2 and dt == pytest.approx(1e-3 / 4)
    assert sim.calls == [1e-3, 5e-4, 2.5e-4]
    assert np.abs(sim.u).max() == pytest.approx(1 + dt * 1.0e4)
    assert sim.t == pytest.approx(dt)  # the rejected attempts did not advance the clock


def test_exhausted_retries_restore_the_state(host):
    sim = HostSim(poison=True)
    u0 = sim.u.copy()
    with pytest.raises(RuntimeError, match="after 2 retries"):
        host.step_adaptive(sim, CFLConstraints(dt_min=1e-8, dt_max=1.0), dt=1e-3, max_retries=2)
    assert np.array_equal(sim.u, u0) and sim.t == 0.0
    assert len(sim.calls) == 3  # the first try and two retries


def test_dt_min_stops_the_halving(host):
    sim = HostSim(poison=True)
    c = CFLConstraints(dt_min=1e-3, dt_max=1.0)
    with pytest.raises(RuntimeError, match="cannot reduce dt further"):
        host.step_adaptive(sim, c, dt=4e-3, max_retries=10)
    assert sim.calls == [4e-3, 2e-3, 1e-3] and sim.t == 0.0


def test_a_throwing_stepper_is_retried(host):
    sim = HostSim()
    seen = []

    def stepper(s, h):
        seen.append(h)
        if h > 0.3:
            raise ValueError("solver diverged")
        s.step(h)

    _, dt, nr = host.step_adaptive(sim, dt=1.0, stepper=stepper, max_retries=3)
    assert seen == [1.0, 0.5, 0.25] and dt == 0.25 and nr == 2


def test_numerical_failure_classification(host):
    sim = HostSim()
    assert host.is_numerical_failure(host.NumericalFailure("x"), sim)
    assert host.is_numerical_failure(ZeroDivisionError(), sim) and host.is_numerical_failure(OverflowError(), sim)
    assert not host.is_numerical_failure(KeyError("k"), sim)          # a bug, from a finite state
    sim.u[0, 0, 0] = math.inf
    assert host.is_numerical_failure(KeyError("k"), sim)              # any throw from a blown-up state
    assert not host.is_numerical_failure(KeyboardInterrupt(), sim)


def test_schedule_validation():
    with pytest.raises(ValueError):
        ts.SubsteppingSchedule(default=0)
    with pytest.raises(ValueError):
        ts.SubsteppingSchedule(microphysics=0)
    assert ts.substeps_for(None, "radiation") == 1
    assert ts.substeps_for(ts.SubsteppingSchedule(microphysics=4), "radiation") == 1


def test_sound_proof_dt_pick_order():
    # adaptive_dt of the sound-proof cores (src/Anelastic.jl:2240-2300):
    # advective, diffusive with ν + max(νt, κt), buoyancy, Coriolis, limits; growth; clamp
    from types import SimpleNamespace
    g = SimpleNamespace(dx=10.0, dy=10.0, dz=10.0)
    p = dict(nu=0.0, C_s=0.17, Pr_t=0.5, f_coriolis=1e-4)
    c = CFLConstraints(dt_max=1e4)
    d2 = 1 / (3 / 100)
    assert ts.mw_dt_pick(1.0, 0.0, 0.0, p, g, c) == (0.5, "advective")
    assert ts.mw_dt_pick(1e-3, 2.0, 0.0, p, g, c) == (0.25 * d2 / 4.0, "diffusive")
    assert ts.mw_dt_pick(1e-3, 0.0, 0.02, p, g, c) == (0.5 / 0.02, "buoyancy")
    assert ts.mw_dt_pick(1e-6, 0.0, 0.0, p, g, c) == (0.5 / 1e-4, "coriolis")
    assert ts.mw_dt_pick(1e-6, 0.0, 0.0, p, g, c, limits=[None, (7.0, "micro")]) == (7.0, "micro")
    assert ts.mw_dt_pick(1e-6, 0.0, 0.0, p, g, c, dt_prev=2.0) == (2.2, "growth")
    assert ts.mw_dt_pick(math.nan, 0.0, 0.0, p, g, c) == (c.dt_min, "nan")
    assert ts.mw_dt_pick(1e-9, 0.0, 0.0, dict(p, f_coriolis=0.0), g, c)[0] == c.dt_max
