"""The static Smagorinsky closure of the step oracle (CPU): on a linear
velocity field the strain stencil is exact, so νt = (C_s Δ)² √((a+c)² + b²)
with Δ = (ΔxΔyΔz)^(1/3); the closure only removes kinetic energy; it
leaves a uniform flow (zero strain) alone."""
import numpy as np


def _linear_case(orc):
    N = (12, 10, 8)
    g = orc.Grid(N, (12 * 4.0, 10 * 3.0, 8 * 2.0), halo=3)
    st = orc.Stepper(g, scheme="centered", bcs=("periodic", "periodic", "wall"),
                     theta_bcs=("periodic", "periodic", "noflux"))
    return g, st


def test_nut_on_linear_field_is_closed_form(orc):
    g, st = _linear_case(orc)
    C_s = 0.17
    st.set_sgs(C_s)
    a, b, c = 0.3, -0.2, 0.5
    sx, sy, sz = g.storage
    H = g.H
    # face positions of u (x faces), v (y faces): u = a·y_c + b·z_c, v = c·x_c
    xf = (np.arange(sx) - H) * g.dx
    xc = xf + 0.5 * g.dx
    yc = (np.arange(sy) - H + 0.5) * g.dy
    zc = (np.arange(sz) - H + 0.5) * g.dz
    u = np.asfortranarray(np.broadcast_to(a * yc[None, :, None] + b * zc[None, None, :], g.storage).copy())
    v = np.asfortranarray(np.broadcast_to(c * xc[:, None, None], g.storage).copy())
    w = g.zeros()
    nut = st.sgs_nut(u, v, w)
    D = (g.dx * g.dy * g.dz) ** (1 / 3)
    want = (C_s * D) ** 2 * np.sqrt((a + c) ** 2 + b ** 2)
    got = nut[g.interior]
    assert np.abs(got - want).max() < 1e-14 * want


def test_closure_dissipates_and_spares_uniform_flow(orc):
    from aeolus_hip import workloads as W
    N = (16, 16, 12)
    g = orc.Grid(N, (16 * 20.0, 16 * 20.0, 12 * 20.0), halo=3)
    rng = np.random.default_rng(4)
    bcs = ("periodic", "periodic", "wall")
    tb = ("periodic", "periodic", "noflux")

    def ke_after(C_s, uniform=False):
        st = orc.Stepper(g, scheme="centered", bcs=bcs, theta_bcs=tb)
        if C_s:
            st.set_sgs(C_s)
        u, v, w, th = (g.zeros() for _ in range(4))
        if uniform:
            u[:] = 2.0
            v[:] = -1.0
        else:
            u[g.interior] = rng.standard_normal(N)
            v[g.interior] = rng.standard_normal(N)
        th[:] = 300.0
        p = g.zeros()
        st.euler(u, v, w, th, p, 1e-9)  # project first
        u0, v0 = u.copy(), v.copy()
        st.euler(u, v, w, th, p, 0.5)
        I = g.interior
        return 0.5 * ((u[I] ** 2).sum() + (v[I] ** 2).sum() + (w[I] ** 2).sum()), u, v, u0, v0
    seed_state = rng.bit_generator.state
    k0 = ke_after(0.0)[0]
    rng.bit_generator.state = seed_state
    k1 = ke_after(0.2)[0]
    assert k1 < k0
    _, u, v, u0, v0 = ke_after(0.2, uniform=True)
    I = g.interior
    assert np.abs(u[I] - 2.0).max() < 1e-12 and np.abs(v[I] + 1.0).max() < 1e-12


def test_scalar_flux_scales_with_inverse_prandtl(orc):
    # the θ change due to the closure is (1/Pr_t)·Δt·∂_j(νt ∂_j θ): halving
    # Pr_t doubles it; the momentum change does not depend on Pr_t
    N = (12, 12, 10)
    g = orc.Grid(N, (12 * 20.0, 12 * 20.0, 10 * 20.0), halo=3)
    rng = np.random.default_rng(8)
    base = [g.zeros() for _ in range(4)]
    base[0][g.interior] = rng.standard_normal(N)
    base[1][g.interior] = rng.standard_normal(N)
    base[3][:] = 300.0
    base[3][g.interior] += rng.standard_normal(N)

    def run(C_s, Pr_t):
        st = orc.Stepper(g, scheme="centered", bcs=("periodic", "periodic", "wall"),
                         theta_bcs=("periodic", "periodic", "noflux"))
        if C_s:
            st.set_sgs(C_s, Pr_t)
        u, v, w, th = (a.copy(order="F") for a in base)
        st.euler(u, v, w, th, g.zeros(), 0.25)
        return u, th
    u0, t0 = run(0.0, 1.0)
    u1, t1 = run(0.2, 1 / 3)
    u2, t2 = run(0.2, 2 / 3)
    I = g.interior
    d1, d2 = t1[I] - t0[I], t2[I] - t0[I]
    assert np.abs(d1).max() > 1e-6
    assert np.abs(d1 - 2 * d2).max() < 1e-9 * np.abs(d1).max()
    assert np.array_equal(u1[I], u2[I]) and np.abs(u1[I] - u0[I]).max() > 1e-6
