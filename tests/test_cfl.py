"""The host half of cfl_diagnostics / adaptive_dt (src/TimeIntegration.jl:
291-466): each process's limit from the state maxima, the binding one (the
SGS limit after the component limits), the growth bound, the clamp and the
NaN fallback."""
import math
from types import SimpleNamespace

import pytest

from aeolus_hip.simulation import CFLConstraints, cfl_limits


def _grid(zf=None):
    return SimpleNamespace(dx=10.0, dy=20.0, dz=5.0, zf=zf)


def _p(**kw):
    p = dict(nu=0.0, C_s=0.0, Pr_t=1 / 3, f_coriolis=0.0)
    p.update(kw)
    return p


def test_each_limit_and_the_binding_one():
    c = CFLConstraints()
    d2 = 1.0 / (1 / 100 + 1 / 400 + 1 / 25)
    d = cfl_limits(2.0, 0.0, 0.0, _p(), _grid(), c)
    assert d["dt_adv"] == 0.5 / 2.0 and d["limiter"] == "advective"
    assert math.isinf(d["dt_diff"]) and math.isinf(d["dt_buoy"]) and math.isinf(d["dt_cor"])
    d = cfl_limits(0.1, 0.0, 0.02, _p(nu=3.0, f_coriolis=1e-4), _grid(), c)
    assert d["dt_diff"] == 0.25 * d2 / 3.0 and d["limiter"] == "diffusive"
    assert d["dt_buoy"] == 0.5 / 0.02 and d["dt_cor"] == 0.5 / 1e-4
    # the SGS limit uses max(νt, κt) with κt = νt/Pr_t and binds last
    d = cfl_limits(0.1, 2.0, 0.0, _p(C_s=0.17, Pr_t=0.5), _grid(), c, limits=[(5.0, "micro")])
    assert d["kappa_max"] == 4.0 and d["dt_sgs"] == 0.25 * d2 / 4.0
    assert d["dt_diff"] == 0.25 * d2 / 2.0
    assert d["limiter"] == "sgs" and d["dt_limits"] == 5.0 and d["limits_limiter"] == "micro"
    # a component limit below every process limit binds, even below dt_sgs
    d = cfl_limits(0.1, 2.0, 0.0, _p(C_s=0.17, Pr_t=0.5), _grid(), c, limits=[(0.01, "micro")])
    assert d["limiter"] == "micro" and d["dt_limiting"] == 0.01
    d = cfl_limits(0.1, 0.0, 0.0, _p(), _grid(), c, limits=[(0.01, "micro"), None, (0.5, "other")])
    assert d["limiter"] == "micro" and d["dt_limiting"] == 0.01
    # a NaN rate poisons the advective limit
    assert math.isnan(cfl_limits(math.nan, 0, 0, _p(), _grid(), c)["dt_adv"])


def test_stretched_grid_thinnest_cell_binds_diffusion():
    import numpy as np
    zf = np.array([0.0, 1.0, 3.0, 7.0])
    d = cfl_limits(0.0, 0.0, 0.0, _p(nu=1.0), _grid(zf), CFLConstraints())
    assert d["dt_diff"] == 0.25 * (1.0 / (1 / 100 + 1 / 400 + 1 / 1.0)) / 1.0


def test_constraints_validation():
    for kw in (dict(C_advective=0), dict(dt_min=2, dt_max=1), dict(growth_max=0.5), dict(C_coriolis=-1)):
        with pytest.raises(ValueError):
            CFLConstraints(**kw)


def test_adaptive_dt_growth_clamp_and_nan():
    from aeolus_hip.simulation import adaptive_dt
    c = CFLConstraints(dt_min=1e-3, dt_max=2.0, growth_max=1.1)
    d = lambda x: dict(dt_limiting=x)
    assert adaptive_dt(None, c, diagnostics=d(1.5)) == 1.5
    assert adaptive_dt(None, c, dt_prev=1.0, diagnostics=d(1.5)) == pytest.approx(1.1)
    assert adaptive_dt(None, c, dt_prev=1.0, diagnostics=d(0.3)) == 0.3      # shrinking is not bounded
    assert adaptive_dt(None, c, diagnostics=d(7.0)) == 2.0
    assert adaptive_dt(None, c, diagnostics=d(1e-6)) == 1e-3
    assert adaptive_dt(None, c, diagnostics=d(math.nan)) == 1e-3
    assert adaptive_dt(None, c, diagnostics=d(math.inf)) == 2.0


def _rate(orc, og, u, v, w):
    import ctypes
    return orc.lib().orc_max_adv_rate(orc.ptr(u), orc.ptr(v), orc.ptr(w), ctypes.byref(og.c))


@pytest.mark.parametrize("stretched", [False, True])
def test_oracle_advective_rate(orc, stretched):
    # _max_advective_rate (src/TimeIntegration.jl:164-182): max over cells of
    # Σ max(|u_face_lo|, |u_face_hi|)/Δx_i, the thinnest cell binding on a stretched grid
    import numpy as np
    zf = np.array([0.0, 1.0, 3.0, 6.0, 10.0, 15.0, 21.0]) if stretched else None
    og = orc.Grid((6, 5, 6), (3.0, 10.0, 21.0), halo=2, zf=zf)
    H = og.H
    dz = np.diff(zf) if stretched else np.full(6, 3.5)
    u, v, w = (og.zeros() for _ in range(3))
    u[...], v[...], w[...] = 2.0, -3.0, 0.5
    want = 2.0 / 0.5 + 3.0 / 2.0 + 0.5 / dz.min()          # uniform flow: closed form
    assert _rate(orc, og, u, v, w) == pytest.approx(want, rel=1e-15)
    rng = np.random.default_rng(4)
    for a in (u, v, w):
        a[...] = rng.standard_normal(a.shape)
    I = slice(H, H + 6), slice(H, H + 5), slice(H, H + 6)
    au = np.maximum(np.abs(u[H:H + 6, I[1], I[2]]), np.abs(u[H + 1:H + 7, I[1], I[2]]))
    av = np.maximum(np.abs(v[I[0], H:H + 5, I[2]]), np.abs(v[I[0], H + 1:H + 6, I[2]]))
    aw = np.maximum(np.abs(w[I[0], I[1], H:H + 6]), np.abs(w[I[0], I[1], H + 1:H + 7]))
    brute = (au / 0.5 + av / 2.0 + aw / dz[None, None, :]).max()
    assert _rate(orc, og, u, v, w) == pytest.approx(brute, rel=1e-14)
    # a NaN anywhere on the faces the interior reads makes the rate NaN
    # (the dt pick then collapses to dt_min), also on the last face of a row
    for a, idx in ((u, (H + 6, H + 1, H + 1)), (v, (H + 2, H, H + 3)), (w, (H, H + 4, H + 6))):
        keep = a[idx]
        a[idx] = np.nan
        assert math.isnan(_rate(orc, og, u, v, w))
        a[idx] = keep
