"""The adaptive step on the GPU: step_guard against max_divergence, numpy
and the oracle; the snapshot's exact rollback; the blow-up case of the
reference's test/ka_step.jl:34-68 (an exhausted retry budget that restores
the state bit for bit, a NaN that collapses the dt pick, a retry that
recovers and equals a direct step at the short dt); the sound-proof cores'
guard (the constraint residual) and dt pick."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _blowup_sim(hip, precision=0):
    # src layout of test/ka_step.jl:35-43: 16x16x8 cells of 10 m, ν = 1e-2,
    # θ₀ = 300, Smagorinsky C_s = 0.17, walls in z; u, v, w = 3·randn; θ = 300 + 0.003 z
    g = hip.Grid((16, 16, 8), (160.0, 160.0, 80.0), halo=3)
    sim = hip.BoussinesqSimulation(g, nu=1e-2, theta0=300.0, C_s=0.17, precision=precision,
                                   bcs=("periodic", "periodic", "wall"),
                                   theta_bcs=("periodic", "periodic", "noflux"))
    rng = np.random.default_rng(3063)
    H, I = g.H, g.interior
    st = [np.zeros(g.storage, order="F") for _ in range(4)]
    for a in st[:3]:
        a[I] = 3 * rng.standard_normal(a[I].shape)
    zc = (np.arange(8) + 0.5) * 10.0
    st[3][:, :, H:H + 8] = 300 + 0.003 * zc[None, None, :]
    sim.set(*st)
    return g, sim


def test_guard_matches_max_divergence_and_speed(hip, orc):
    g, sim = _blowup_sim(hip)
    # an unprojected state: the divergence is O(1), so the comparison is not at roundoff
    gd = hip.step_guard(sim)
    assert gd["finite"]
    assert gd["divergence"] == sim.max_divergence()
    st = sim.state()
    I = g.interior
    assert gd["speed"] == max(np.abs(st[n][I]).max() for n in ("u", "v", "w"))
    og = orc.Grid((16, 16, 8), (160.0, 160.0, 80.0), halo=3)
    ref = orc.max_divergence(og, st["u"], st["v"], st["w"])
    assert abs(gd["divergence"] - ref) <= 1e-14 * ref
    # a NaN in p and an Inf in θ are seen; the velocity maxima stay as they were
    for name, val in (("p", np.nan), ("theta", np.inf)):
        a = st[name].copy()
        a[g.H + 5, g.H + 2, g.H + 7] = val
        getattr(sim, name).upload(a)
        bad = hip.step_guard(sim)
        assert not bad["finite"] and bad["speed"] == gd["speed"]
        getattr(sim, name).upload(st[name])
    assert hip.step_guard(sim)["finite"]


@pytest.mark.parametrize("prec", [0, 1])
def test_guard_on_a_stretched_grid(hip, orc, prec):
    from aeolus_hip import workloads as W
    zf, Lz = W.dcbl_faces(20)
    g = hip.Grid((24, 16, 20), (480.0, 320.0, Lz), halo=3, zf=zf)
    sim = hip.BoussinesqSimulation(g, precision=prec, **W.DCBL_PHYS)
    rng = np.random.default_rng(8)
    st = [np.zeros(g.storage, order="F") for _ in range(4)]
    for a in st[:3]:
        a[g.interior] = rng.standard_normal(a[g.interior].shape)
    sim.set(*st)
    gd = hip.step_guard(sim)
    assert gd["finite"] and gd["divergence"] == sim.max_divergence()
    og = orc.Grid((24, 16, 20), (480.0, 320.0, Lz), halo=3, zf=zf)
    s = sim.state()
    ref = orc.max_divergence(og, *(s[n].astype(np.float64) for n in ("u", "v", "w")))
    assert abs(gd["divergence"] - ref) <= (1e-14 if prec == 0 else 1e-5) * ref


def test_snapshot_roundtrip_is_exact(hip):
    g, sim = _blowup_sim(hip)
    sim.step(0.01)
    snap = hip.state_snapshot(sim)
    assert snap is hip.state_snapshot(sim)  # kept on the simulation
    with pytest.raises(ValueError):
        hip.restore_state(sim, snap)        # never saved
    before = sim.state()
    hip.save_state(snap, sim)
    sim.step(0.02)
    assert sim.t == pytest.approx(0.03)
    hip.restore_state(sim, snap)
    after = sim.state()
    assert sim.t == 0.01
    for n in before:
        assert np.array_equal(after[n], before[n]), n


def test_blowup_restore_nan_pick_and_recovering_retry(hip):
    g, sim = _blowup_sim(hip)
    c = hip.CFLConstraints(dt_max=100.0, dt_min=1e-6)
    sim.step(0.01)  # a projected state
    s0, t0 = sim.state(), sim.t
    # dt = 1e4 runs away; every retry (dt halves) exceeds the speed cap
    with pytest.raises(RuntimeError, match="after 2 retries"):
        hip.step_adaptive(sim, c, dt=1e4, max_retries=2, max_velocity=50.0)
    s1 = sim.state()
    for n in s0:  # copied back, not recomputed: exact, halos included
        assert np.array_equal(s1[n], s0[n]), n
    assert sim.t == t0
    # a NaN velocity: the dt pick collapses to dt_min and the guard sees it
    u = s0["u"].copy()
    u[g.H + 1, g.H + 1, g.H + 1] = np.nan
    sim.u.upload(u)
    assert hip.adaptive_dt(sim, c) == c.dt_min
    assert not hip.step_guard(sim)["finite"]
    sim.u.upload(s0["u"])
    # a retry that recovers equals a direct step at the short dt
    _, gb = _blowup_sim(hip)
    for n in ("u", "v", "w", "theta", "p"):
        getattr(gb, n).upload(s0[n])
    _, dt, nr = hip.step_adaptive(sim, c, dt=40.0, max_retries=12, max_velocity=50.0)
    assert nr > 0 and dt == 40.0 / 2 ** nr
    gb.step(dt)
    a, b = sim.state(), gb.state()
    I = g.interior
    for n in ("u", "v", "w"):
        assert np.array_equal(a[n][I], b[n][I]), n  # the guard filled a's velocity halos
    for n in ("theta", "p"):
        assert np.array_equal(a[n], b[n]), n
    assert sim.t == pytest.approx(t0 + dt)


def test_adaptive_run_with_rk3_and_fast_processes(hip):
    g, sim = _blowup_sim(hip)
    c = hip.CFLConstraints(dt_max=5.0)
    sim.step(0.01)
    calls, dt = [], None
    for _ in range(3):
        _, dt, nr = hip.step_adaptive(sim, c, dt_prev=dt, stepper=lambda s, h: s.step_rk3(h),
                                      fast_processes=dict(microphysics=lambda s, h: calls.append(h)),
                                      schedule=hip.SubsteppingSchedule(microphysics=2))
        assert nr == 0
    assert len(calls) == 6 and calls[-1] == pytest.approx(dt / 2)
    gd = hip.step_guard(sim)
    assert gd["finite"] and gd["divergence"] < 1e-10


def _mw(hip, orc, weighting, C_s=0.0):
    from aeolus_hip.anelastic import AnelasticSimulation, PseudoIncompressibleSimulation, isothermal_reference
    N, H, dz = (16, 8, 20), 1, 100.0
    extent = (N[0] * 200.0, N[1] * 200.0, N[2] * dz)
    g, og = hip.Grid(N, extent, halo=H), orc.Grid(N, extent, halo=H)
    ref = isothermal_reference(N[2], dz, H)
    rng = np.random.default_rng(44)
    st = og.storage
    th = np.zeros(st)
    th[:, :, H:H + N[2]] = ref.theta[None, None, :]
    I = og.interior
    th[I] += 0.5 * rng.standard_normal(th[I].shape)
    u, v, w = (np.zeros(st) for _ in range(3))
    for a in (u, v, w):
        a[I] += rng.standard_normal(a[I].shape)
    u, v, w, th = (np.asfortranarray(a) for a in (u, v, w, th))
    cls = AnelasticSimulation if weighting == "anelastic" else PseudoIncompressibleSimulation
    sim = cls(g, ref, nu=0.5, flux_bot=0.05, C_s=C_s, f_coriolis=1e-4)
    sim.set(u, v, w, th)
    return g, og, ref, sim, (u, v, w, th)


@pytest.mark.parametrize("weighting", ["anelastic", "pseudo_incompressible"])
def test_sound_proof_guard_is_the_constraint_residual(hip, orc, weighting):
    g, og, ref, sim, (u, v, w, th) = _mw(hip, orc, weighting)
    gd = hip.step_guard(sim)
    bcs = ("periodic", "periodic", "wall")
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(og, a, loc, bcs)
    _, _, wc_pad, wf_pad = ref.weights(weighting)
    d = orc.weighted_divergence(og, u, v, w, np.append(wc_pad, wc_pad[-1]), np.append(wf_pad, wf_pad[-1]),
                                normalize=True)
    res = np.abs(d[og.interior]).max()
    assert gd["finite"] and abs(gd["divergence"] - res) <= 1e-13 * res
    assert gd["speed"] == max(np.abs(a[og.interior]).max() for a in (u, v, w))
    # after a step the residual is at the projection's roundoff
    sim.step(3.0)
    assert hip.step_guard(sim)["divergence"] < 1e-12


@pytest.mark.parametrize("weighting", ["anelastic", "pseudo_incompressible"])
def test_sound_proof_dt_pick(hip, orc, weighting):
    from aeolus_hip.timestepping import mw_cfl_maxima, mw_dt_pick
    g, og, ref, sim, (u, v, w, th) = _mw(hip, orc, weighting, C_s=0.17)
    rate, nut, N = mw_cfl_maxima(sim)
    bcs = ("periodic", "periodic", "wall")
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(og, a, loc, bcs)
    r = orc.lib().orc_max_adv_rate(orc.ptr(u), orc.ptr(v), orc.ptr(w), ctypes.byref(og.c))
    assert abs(rate - r) <= 1e-14 * r
    # νt = (C_s Δ)²|S| is the same pointwise closure on either core (regular grid)
    ost = orc.Stepper(og, bcs=bcs)
    ost.set_sgs(0.17)
    nu_ref = ost.sgs_nut(u, v, w)[og.interior].max()
    assert nu_ref > 0 and abs(nut - nu_ref) <= 1e-13 * nu_ref
    m = th[og.interior].mean(axis=(0, 1))
    n2 = max(0.0, max((ref.g / ((m[k - 1] + m[k]) / 2)) * (m[k] - m[k - 1]) / g.dz for k in range(1, g.Nz)))
    assert abs(N - np.sqrt(n2)) <= 1e-12 * max(np.sqrt(n2), 1e-30)
    c = hip.CFLConstraints(dt_max=1e4)
    want = mw_dt_pick(r, nut, N, sim.parameters(), g, c, dt_prev=1.0)[0]
    assert hip.adaptive_dt(sim, c, dt_prev=1.0) == pytest.approx(want, rel=1e-12)
    # the adaptive step on the sound-proof core: a dt far above the pick is halved until accepted
    _, dt, nr = hip.step_adaptive(sim, c, dt=8 * want, max_retries=8, max_velocity=20.0)
    gd = hip.step_guard(sim)
    assert gd["finite"] and gd["speed"] <= 20.0 and dt == 8 * want / 2 ** nr
