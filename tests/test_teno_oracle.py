"""TENO5/TENO8 and the variance-conserving operator in the oracle
(src/AdvectionSchemes.jl:391-495, 1407-1446), pinned by the reference's
tests (test/advection.jl:52-90, 121-141) and by closed forms:

- TENO5 reproduces a degree-4 polynomial's face value to 1e-13 (the linear
  weights verbatim); TENO8 a degree-7 polynomial's (the 8th-order upwind
  reconstruction when all six candidates survive).
- observed order > 4.8 (TENO5) and > 7 (TENO8) on cell averages of sin.
- at a jump the cut-off removes both stencils that cross it and TENO5
  returns the smooth candidate's value q0 to an ulp, where WENO-Z5 keeps a
  ~1e-5 share of the others: the discrete selection, checked by hand.
- the square wave of test/advection.jl:70-90 after one period of SSP-RK3:
  TV ratio < 1.02, over- and undershoot < 5e-3, mass < 1e-12.
- the skew form conserves Σϕ² for a divergent velocity field, where the
  centred flux form does not.
"""
import math

import numpy as np
import pytest


def _avgs(F, x0, h, off0, n):
    return [(F(x0 - h / 2 + (off0 + j) * h + h / 2) - F(x0 - h / 2 + (off0 + j) * h - h / 2)) / h for j in range(n)]


def _teno(orc, n, v, C_T=1e-5):
    return orc.lib().orc_teno5_left(*v, C_T) if n == 5 else orc.lib().orc_teno8_left(*v, C_T)


def test_teno5_degree4_exact(orc):
    c = [1.0, 2.0, -0.5, 0.3, -0.1]
    p4 = lambda x: sum(c[m] * x ** m for m in range(5))
    P4 = lambda x: sum(c[m] * x ** (m + 1) / (m + 1) for m in range(5))
    assert abs(_teno(orc, 5, _avgs(P4, 0.7, 0.3, -2, 5)) - p4(0.7)) < 1e-13


def test_teno8_degree7_exact(orc):
    c = [0.5, -1.0, 0.7, 0.2, -0.3, 0.1, 0.05, -0.02]
    p7 = lambda x: sum(c[m] * x ** m for m in range(8))
    P7 = lambda x: sum(c[m] * x ** (m + 1) / (m + 1) for m in range(8))
    assert abs(_teno(orc, 8, _avgs(P7, 0.4, 0.1, -4, 8)) - p7(0.4)) < 1e-12


@pytest.mark.parametrize("n,off0,order", [(5, -2, 4.8), (8, -4, 7.0)])
def test_teno_observed_order(orc, n, off0, order):
    x0 = 0.3
    e = [abs(_teno(orc, n, _avgs(lambda x: -math.cos(x), x0, h, off0, n)) - math.sin(x0)) for h in (0.2, 0.1)]
    assert math.log(e[0] / e[1]) / math.log(2.0) > order


def test_teno5_cutoff_at_a_step(orc):
    # a smooth left stencil, a jump between cells 3 and 4: by hand β0 = 0.0025,
    # β1 ≈ 1.42, β2 ≈ 3.33, so g1/Σg ≈ 2.5e-16 and g2/Σg are below C_T and
    # only the left candidate q0 = (2·0.1 − 7·0.05 + 0)/6 survives
    v = (0.1, 0.05, 0.0, 1.0, 1.0)
    q0 = ((2 * v[0] - 7 * v[1]) + 11 * v[2]) / 6
    assert abs(_teno(orc, 5, v) - q0) <= 2 * np.spacing(abs(q0))
    # WENO-Z5 keeps a ω1 ≈ 2e-5 share of the stencil across the jump
    assert abs(orc.lib().orc_weno_z5_left(*v, 2) - q0) > 1e-6
    # with the cut-off at 0 every candidate survives: the linear 5th-order value
    lin = (2 * v[0] - 13 * v[1] + 47 * v[2] + 27 * v[3] - 3 * v[4]) / 60
    assert abs(_teno(orc, 5, v, C_T=1e-300) - lin) < 1e-15


def _rk3_square_wave(orc, scheme):
    og = orc.Grid((64, 4, 4), (1.0, 1 / 16, 1 / 16), halo=3)
    H = og.H
    u, v, w = og.zeros(), og.zeros(), og.zeros()
    u[...] = 1.0
    bcs = ("periodic",) * 3
    dt = 0.4 / 64
    n = round(1.0 / dt)
    x = (np.arange(64) + 0.5) / 64
    phi = og.zeros()
    phi[H:H + 64, H:H + 4, H:H + 4] = ((x >= 0.25) & (x <= 0.5)).astype(float)[:, None, None]
    I = og.interior
    a0 = phi[H:H + 64, H, H].copy()
    m0 = phi[I].sum()

    def stage(ph0, ph, a, b):
        orc.fill_halos(og, ph, orc.CC, bcs)
        t = og.zeros()
        orc.advect_scalar(og, t, u, v, w, ph, scheme)
        ph[I] = a * ph0[I] + b * (ph[I] + dt * t[I])

    for _ in range(n):
        ph0 = phi.copy(order="F")
        stage(ph0, phi, 0.0, 1.0)
        stage(ph0, phi, 0.75, 0.25)
        stage(ph0, phi, 1 / 3, 2 / 3)
    a = phi[H:H + 64, H, H]
    tv = lambda z: np.abs(np.diff(np.append(z, z[0]))).sum()
    return a.max() - 1, -a.min(), tv(a) / tv(a0), (phi[I].sum() - m0) / m0


def test_teno5_square_wave(orc):
    over, under, tv, mass = _rk3_square_wave(orc, "teno5")
    assert tv < 1.02 and over < 5e-3 and under < 5e-3 and abs(mass) < 1e-12
    over, under, tv, mass = _rk3_square_wave(orc, "centered")
    assert tv > 2.0 and abs(mass) < 1e-12


def test_variance_conserving(orc):
    og = orc.Grid((8, 8, 8), (1.0, 1.0, 1.0), halo=3)
    rng = np.random.default_rng(2718)
    I = og.interior
    arrs = []
    for loc in (orc.FX, orc.FY, orc.FZ, orc.CC):
        a = og.zeros()
        a[I] = 1 + rng.standard_normal(a[I].shape)
        orc.fill_halos(og, a, loc, ("periodic",) * 3)
        arrs.append(a)
    u, v, w, phi = arrs
    t = orc.advect_variance_conserving(og, og.zeros(), u, v, w, phi)
    dV = (1 / 8) ** 3
    vs, ms = (phi[I] * t[I]).sum() * dV, t[I].sum() * dV
    tc = orc.advect_scalar(og, og.zeros(), u, v, w, phi, "centered")
    vc = (phi[I] * tc[I]).sum() * dV
    assert abs(vs) < 1e-12
    assert abs(vc) > 1e-3      # the centred flux form does not conserve Σϕ² here
    assert abs(ms) > 1e-3      # nor does the skew form conserve Σϕ when ∇·u ≠ 0
