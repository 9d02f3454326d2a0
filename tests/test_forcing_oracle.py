"""Coriolis and Rayleigh-sponge terms of the step oracle (CPU): a
horizontally uniform profile under the sponge decays by exactly the sin²
rate of sponge_profile (src/FiniteVolume.jl:524-566) and is untouched below
z_start; a uniform flow on the f-plane turns at the inertial frequency."""
import numpy as np

from helpers import rel_err


def _profile_case(orc, stretched):
    from aeolus_hip import workloads as W
    N = (8, 6, 20)
    if stretched:
        zf, Lz = W.dcbl_faces(N[2])
        g = orc.Grid(N, (160.0, 120.0, Lz), halo=3, zf=zf)
        zc_ref = (zf[:-1] + zf[1:]) / 2
        zw_ref = zf[:-1]
    else:
        Lz = 2000.0
        g = orc.Grid(N, (160.0, 120.0, Lz), halo=3)
        dz = Lz / N[2]
        zc_ref = (np.arange(N[2]) + 0.5) * dz
        zw_ref = np.arange(N[2]) * dz
    return g, Lz, zc_ref, zw_ref


def _sin2(z, alpha, z0, z1):
    r = np.where(z >= z0, alpha * np.sin(0.5 * np.pi * (z - z0) / (z1 - z0)) ** 2, 0.0)
    return r


def test_sponge_decays_profile_at_sin2_rate(orc):
    for stretched in (False, True):
        g, Lz, zc, zw = _profile_case(orc, stretched)
        st = orc.Stepper(g, scheme="centered", bcs=("periodic", "periodic", "wall"),
                         theta_bcs=("periodic", "periodic", "noflux"))
        alpha, z0 = 0.05, 0.5 * Lz
        tu = np.linspace(0.0, 1.0, g.Nz)
        st.set_forcing(0.0, alpha, z0, Lz, zc, zw, (tu, None, None))
        H = g.H
        u, v, w, th = (g.zeros() for _ in range(4))
        prof = 2.0 + np.cos(np.arange(g.Nz))
        u[:, :, H:H + g.Nz] = prof[None, None, :]
        th[:] = 300.0
        p = g.zeros()
        dt = 0.7
        st.euler(u, v, w, th, p, dt)
        lam = _sin2(zc, alpha, z0, Lz)
        want = prof - dt * lam * (prof - tu)
        got = u[g.interior]
        assert rel_err(got, np.broadcast_to(want, got.shape)) < 1e-13
        below = zc < z0
        assert np.array_equal(got[:, :, below], np.broadcast_to(prof[below], got[:, :, below].shape))
        assert np.abs(v[g.interior]).max() == 0


def test_inertial_oscillation(orc):
    g = orc.Grid((8, 8, 8), (1.0, 1.0, 1.0), halo=3)
    st = orc.Stepper(g, scheme="weno5")
    f = 1e-2
    st.set_forcing(f)
    U, V = 3.0, -1.0
    u, v, w, th = (g.zeros() for _ in range(4))
    u[:] = U
    v[:] = V
    p = g.zeros()
    dt, n = 1.0, 100
    for _ in range(n):
        st.rk3(u, v, w, th, p, dt)
    t = n * dt
    ue = U * np.cos(f * t) + V * np.sin(f * t)
    ve = -U * np.sin(f * t) + V * np.cos(f * t)
    I = g.interior
    assert abs(u[I] - ue).max() < 2e-7 and abs(v[I] - ve).max() < 2e-7
    # on this linear system SSP-RK3 is the cubic Taylor polynomial of the
    # rotation: the discrete solution is R^n (U, V) up to round-off
    A = f * dt * np.array([[0.0, 1.0], [-1.0, 0.0]])
    R = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6
    ud, vd = np.linalg.matrix_power(R, n) @ np.array([U, V])
    assert abs(u[I] - ud).max() < 1e-13 and abs(v[I] - vd).max() < 1e-13
    # the flow stays uniform: the 4-point averages see one value
    assert np.ptp(u[I]) < 1e-13 and np.ptp(v[I]) < 1e-13
