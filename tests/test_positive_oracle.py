"""The oracle's advect_positive! restatement (src/AdvectionSchemes.jl:
858-1160), pinned by the reference's own tests and by closed forms.

- The narrow Gaussian at CFL 0.95 (test/advection.jl:92-119): 20 forward
  Euler steps of unlimited WENO-Z5 go below −0.1; the limited ones stay
  ≥ −1e-14·max with mass conserved to 1e-12.
- No outflow anywhere leaves every ratio at 1, with ϕ > 0 and with ϕ = 0
  (test/advection_branchfree.jl:92-96).
- One donor cell in a zero field, centred scheme, uniform U > 0: by hand, the
  donor's west neighbour has Q = 0 and antidiffusive outflow Uq/2, so R = 0
  there and the donor's west-face flux is limited to its upwind value 0; the
  east face keeps Uq/2. A swapped donor index or a dropped low-order flux
  changes these numbers.
- A smooth positive field: the limiter is inactive (R = 1) and the tendency is
  the unlimited flux-form WENO advection of the same oracle.
"""
import numpy as np
import pytest


def _periodic_velocity(orc, og, U):
    u, v, w = og.zeros(), og.zeros(), og.zeros()
    u[...], v[...], w[...] = U
    return u, v, w


def test_narrow_gaussian_stays_positive(orc):
    og = orc.Grid((16, 16, 16), (1.0, 1.0, 1.0), halo=3)
    h = 1.0 / 16
    u, v, w = _periodic_velocity(orc, og, (1.0, 0.6, 0.3))
    dt = 0.95 / (1.0 / h + 0.6 / h + 0.3 / h)
    s = 1.2 * h
    x = (np.arange(16) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    gauss = np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) / (2 * s * s))
    I = og.interior
    bcs = ("periodic",) * 3

    def euler(limited):
        phi = og.zeros()
        phi[I] = gauss
        m0 = phi[I].sum()
        for _ in range(20):
            orc.fill_halos(og, phi, orc.CC, bcs)
            t = og.zeros()
            if limited:
                orc.advect_positive(og, t, u, v, w, phi, dt, scheme="weno_z5")
            else:
                orc.advect_scalar(og, t, u, v, w, phi, "weno_z5")
            phi[I] += dt * t[I]
        return phi[I].min(), phi[I].max(), (phi[I].sum() - m0) / m0

    rmin, _, _ = euler(False)
    lmin, lmax, lmass = euler(True)
    assert rmin < -0.1
    assert lmin >= -1e-14 * lmax
    assert abs(lmass) < 1e-12


def test_no_outflow_gives_unit_ratio(orc):
    og = orc.Grid((6, 5, 4), (1.0, 1.0, 1.0), halo=3)
    u, v, w = og.zeros(), og.zeros(), og.zeros()
    for val in (1.5, 0.0):
        phi = og.zeros()
        phi[...] = val
        _, R = orc.advect_positive(og, og.zeros(), u, v, w, phi, 1.0, scheme="weno_z5", ratio=True)
        assert np.all(R[og.interior] == 1.0)


@pytest.mark.parametrize("bcs", [None, ("periodic",) * 3])
def test_single_donor_closed_form(orc, bcs):
    og = orc.Grid((8, 4, 4), (8.0, 4.0, 4.0), halo=3)  # Δ = 1
    U, q, dt = 0.5, 2.0, 0.4
    u, v, w = _periodic_velocity(orc, og, (U, 0.0, 0.0))
    phi = og.zeros()
    H = og.H
    d = H + 3                         # donor cell (0-based storage index)
    phi[d, :, :] = q
    t, R = orc.advect_positive(og, og.zeros(), u, v, w, phi, dt, scheme="centered", bcs=bcs, ratio=True)
    J, K = slice(H, H + 4), slice(H, H + 4)
    assert np.all(R[d - 1, J, K] == 0.0)          # Q = 0, outflow Uq/2
    assert np.all(R[d, J, K] == 1.0)              # no antidiffusive outflow
    assert np.all(R[d + 1, J, K] == 1.0)          # Q/P = 2, capped at 1
    assert np.all(t[d - 1, J, K] == 0.0)          # unlimited centred: −Uq/2
    assert np.all(t[d, J, K] == -U * q / 2)
    assert np.all(t[d + 1, J, K] == U * q / 2)
    others = [i for i in range(H, H + 8) if i not in (d - 1, d, d + 1)]
    assert np.all(t[others][:, J, K] == 0.0)
    # unlimited, the west neighbour would go negative
    tc = orc.advect_scalar(og, og.zeros(), u, v, w, phi, "centered")
    assert tc[d - 1, H, H] == -U * q / 2


@pytest.mark.parametrize("scheme", ["weno_z5", "weno_z7", "weno5", "centered"])
def test_smooth_field_is_unlimited(orc, scheme):
    H = 4 if scheme == "weno_z7" else 3
    og = orc.Grid((12, 10, 8), (1.0, 1.0, 1.0), halo=H)
    rng = np.random.default_rng(9)
    u, v, w = (og.zeros() for _ in range(3))
    for a in (u, v, w):
        a[...] = 0.3 + 0.05 * rng.standard_normal(a.shape)
    bcs = ("periodic",) * 3
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(og, a, loc, bcs)
    x = (np.arange(og.storage[0]) - H + 0.5) / 12
    phi = og.zeros()
    phi[...] = 3.0 + np.sin(2 * np.pi * x)[:, None, None]
    orc.fill_halos(og, phi, orc.CC, bcs)
    t, R = orc.advect_positive(og, og.zeros(), u, v, w, phi, 0.01, scheme=scheme, ratio=True)
    assert np.all(R[og.interior] == 1.0)
    ref = orc.advect_scalar(og, og.zeros(), u, v, w, phi, "weno_js5" if scheme == "weno5" else scheme)
    I = og.interior
    assert np.abs(t[I] - ref[I]).max() <= 1e-13 * np.abs(ref[I]).max()


def test_random_positive_field_walls_and_accumulate(orc):
    # walls in z (w = 0 on the wall faces), random positive ϕ, random flow; one
    # Euler update at first-order CFL 0.9 stays non-negative, the total is kept
    og = orc.Grid((10, 8, 12), (10.0, 8.0, 12.0), halo=3)
    rng = np.random.default_rng(31)
    u, v, w = (og.zeros() for _ in range(3))
    for a in (u, v, w):
        a[...] = rng.standard_normal(a.shape)
    bcs = ("periodic", "periodic", "wall")
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(og, a, loc, bcs)
    phi = og.zeros()
    phi[...] = rng.uniform(0, 1, phi.shape) ** 4
    orc.fill_halos(og, phi, orc.CC, ("periodic", "periodic", "noflux"))
    I = og.interior
    H = og.H
    rate = (np.maximum(np.abs(u[H:H + 10, H:H + 8, H:H + 12]), np.abs(u[H + 1:H + 11, H:H + 8, H:H + 12]))
            + np.maximum(np.abs(v[H:H + 10, H:H + 8, H:H + 12]), np.abs(v[H:H + 10, H + 1:H + 9, H:H + 12]))
            + np.maximum(np.abs(w[H:H + 10, H:H + 8, H:H + 12]), np.abs(w[H:H + 10, H:H + 8, H + 1:H + 13])))
    dt = 0.9 / rate.max() / 2  # Σ|u_d|Δt/Δx_d ≤ 1 with both faces of a cell outflowing
    for bc in (None, ("periodic", "periodic", "noflux")):
        t = orc.advect_positive(og, og.zeros(), u, v, w, phi, dt, scheme="weno_z5", bcs=bc)
        new = phi[I] + dt * t[I]
        assert new.min() >= -1e-15 * phi[I].max()
        assert abs(t[I].sum()) <= 1e-12 * np.abs(t[I]).sum()
        raw = orc.advect_scalar(og, og.zeros(), u, v, w, phi, "weno_z5")
        assert (phi[I] + dt * raw[I]).min() < 0  # the limiter had work to do
    # accumulate = False writes −∇·F; accumulate adds it to what tend holds
    base = og.zeros()
    base[...] = rng.standard_normal(base.shape)
    t_set = orc.advect_positive(og, base.copy(order="F"), u, v, w, phi, dt, accumulate=False)
    t_add = orc.advect_positive(og, base.copy(order="F"), u, v, w, phi, dt)
    assert np.allclose(t_add[I] - base[I], t_set[I], rtol=0, atol=1e-13 * np.abs(t_set[I]).max())
    assert np.array_equal(t_set[0], base[0])  # halo untouched
