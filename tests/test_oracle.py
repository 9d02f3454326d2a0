"""Pin the CPU oracle: reference known answers, closed forms, brute force.

Each check is chosen so a plausible restatement mistake (a wrong indicator
coefficient, a transposed stencil, a wrong metric, a sign) fails it. These
run on the CPU only (no GPU).
"""
import math
import os

import numpy as np
import pytest

from helpers import cell_avgs_sin, rand_storage, rel_err


# ---------------------------------------------------------------------------
# reconstructions
# ---------------------------------------------------------------------------
def order(f, n, off0, hs=(0.2, 0.1), x0=0.3):
    e = [abs(f(*cell_avgs_sin(x0, h, off0, n)) - math.sin(x0)) for h in hs]
    return math.log(e[0] / e[1]) / math.log(hs[0] / hs[1])


def test_weno_z5_design_order(orc):
    # test/advection.jl:53 (observed order > 4.8)
    L = orc.lib()
    assert order(lambda *v: L.orc_weno_z5_left(*v, 2), 5, -2) > 4.8
    assert order(lambda *v: L.orc_weno_z5_left(*v, 1), 5, -2) > 4.8


def test_weno_z7_design_order(orc):
    L = orc.lib()
    assert order(lambda *v: L.orc_weno_z7_left(*v, 2), 7, -3, hs=(0.4, 0.2)) > 6.5


def test_weno_z_beats_js_at_critical_point(orc):
    # test/advection.jl:64-67
    L = orc.lib()
    h = 0.1
    x0 = math.pi / 2 + h / 2
    v = [(math.cos(x0 - h / 2 + (j - 3) * h - h / 2) - math.cos(x0 - h / 2 + (j - 3) * h + h / 2)) / h
         for j in range(1, 6)]
    ez = abs(L.orc_weno_z5_left(*v, 2) - math.sin(x0))
    ej = abs(L.orc_weno_js5_left(*v, 1e-6) - math.sin(x0))
    assert ez < ej


def _js_textbook(v, eps=1e-6):
    """WENO-JS with the textbook indicators 13/12(v1-2v2+v3)^2 + 1/4(...)^2 (the
    expanded form the reference's difference basis reproduces identically)."""
    v1, v2, v3, v4, v5 = v
    b0 = 13 / 12 * (v1 - 2 * v2 + v3) ** 2 + 1 / 4 * (v1 - 4 * v2 + 3 * v3) ** 2
    b1 = 13 / 12 * (v2 - 2 * v3 + v4) ** 2 + 1 / 4 * (v2 - v4) ** 2
    b2 = 13 / 12 * (v3 - 2 * v4 + v5) ** 2 + 1 / 4 * (3 * v3 - 4 * v4 + v5) ** 2
    q = [(2 * v1 - 7 * v2 + 11 * v3) / 6, (-v2 + 5 * v3 + 2 * v4) / 6, (2 * v3 + 5 * v4 - v5) / 6]
    a = [0.1 / (b0 + eps) ** 2, 0.6 / (b1 + eps) ** 2, 0.3 / (b2 + eps) ** 2]
    return sum(x * y for x, y in zip(a, q)) / sum(a)


def test_weno_js_indicators_equal_textbook(orc):
    L = orc.lib()
    rng = np.random.default_rng(7)
    for _ in range(2000):
        v = list(rng.normal(size=5))
        ref = _js_textbook(v)
        assert abs(L.orc_weno_js5_left(*v, 1e-6) - ref) <= 1e-12 * max(1.0, abs(ref))
        assert abs(L.orc_weno5_classic_left(*v) - ref) <= 1e-12 * max(1.0, abs(ref))


def test_weno_z5_smooth_limit_is_linear_scheme(orc):
    # on a linear profile every indicator is equal, τ = 0 and the weights are
    # the ideal (1, 6, 3)/10: the 5th-order linear upwind value, exact for lines
    L = orc.lib()
    for s in (0.5, -2.0, 3.25):
        v = [1.0 + s * m for m in range(5)]
        face = 1.0 + s * 2.5  # cell averages of a line: the value at the interface i+1/2
        assert abs(L.orc_weno_z5_left(*v, 2) - face) < 1e-13
        assert abs(L.orc_weno_z7_left(*[1.0 + s * m for m in range(7)], 2) - (1.0 + s * 3.5)) < 1e-13


def test_weno_z5_degree4_linear_weights(orc):
    # with the ideal weights the three candidates reproduce the degree-4
    # polynomial's face value exactly; WENO-Z approaches them on smooth data
    L = orc.lib()
    c = [1.0, 2.0, -0.5, 0.3, -0.1]
    P4 = lambda x: sum(c[m] * x ** (m + 1) / (m + 1) for m in range(5))
    p4 = lambda x: sum(c[m] * x ** m for m in range(5))
    h, x0 = 0.01, 0.7
    v = [(P4(x0 - h / 2 + (j - 2) * h + h / 2) - P4(x0 - h / 2 + (j - 2) * h - h / 2)) / h for j in range(5)]
    assert abs(L.orc_weno_z5_left(*v, 2) - p4(x0)) < 1e-10


# ---------------------------------------------------------------------------
# halo fills
# ---------------------------------------------------------------------------
def test_halo_periodic_mirror_wall(orc):
    g = orc.Grid((6, 5, 7), (1.0, 1.0, 1.0), halo=3)
    a = rand_storage(g.storage, 1)
    b = a.copy(order="F")
    orc.fill_halos(g, b, orc.FZ, ("periodic", "noflux", "wall"))
    H = 3
    # periodic x: b[h] = b[N + h]
    assert np.array_equal(b[0:H, H:-H, H + 1:-H], b[6:6 + H, H:-H, H + 1:-H])
    # mirror y: b[:, H-1-h] = b[:, H+h]
    for h in range(H):
        assert np.array_equal(b[H:-H, H - 1 - h, H + 1:-H], b[H:-H, H + h, H + 1:-H])
    # wall for w: the wall faces are zero, the halo antisymmetric about them
    assert np.all(b[:, :, H] == 0) and np.all(b[:, :, H + 7] == 0)
    for h in range(1, H + 1):
        assert np.array_equal(b[:, :, H - h], -b[:, :, H + h])
    for h in range(1, H):
        assert np.array_equal(b[:, :, H + 7 + h], -b[:, :, H + 7 - h])
    # interior other than the wall plane untouched
    assert np.array_equal(b[H:-H, H:-H, H + 1:H + 7], a[H:-H, H:-H, H + 1:H + 7])


# ---------------------------------------------------------------------------
# advection operators
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", ["centered", "upwind", "weno5", "weno_js5", "weno_z5", "weno_z7"])
def test_flux_form_conserves_mass(orc, scheme):
    # test/advection.jl:145-146, test/runtests.jl:482-499: |Σ tend| < 1e-12
    H = 4 if scheme == "weno_z7" else 3
    g = orc.Grid((12, 10, 8), (1.0, 1.0, 1.0), halo=H)
    u, v, w, phi = (rand_storage(g.storage, s) for s in (1, 2, 3, 4))
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ), (phi, orc.CC)):
        orc.fill_halos(g, a, loc, ("periodic",) * 3)
    t = g.zeros()
    orc.advect_scalar(g, t, u, v, w, phi, scheme)
    assert abs(t[g.interior].sum()) < 1e-12 * np.abs(t[g.interior]).max() * t[g.interior].size
    tu, tv, tw = g.zeros(), g.zeros(), g.zeros()
    orc.advect_momentum(g, tu, tv, tw, u, v, w, scheme)
    for a in (tu, tv, tw):
        assert abs(a[g.interior].sum()) < 1e-12 * np.abs(a[g.interior]).max() * a[g.interior].size


def test_constant_scalar_gives_minus_c_div(orc):
    # every reconstruction of a constant is the constant: tend = -c ∇·u
    g = orc.Grid((10, 9, 8), (2.0, 1.0, 1.5), halo=3)
    u, v, w = (rand_storage(g.storage, s) for s in (5, 6, 7))
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(g, a, loc, ("periodic",) * 3)
    c = 3.7
    phi = np.full(g.storage, c, order="F")
    div = orc.divergence_interior(g, u, v, w)
    for scheme in ("weno_z5", "weno_js5", "weno5", "centered"):
        t = g.zeros()
        orc.advect_scalar(g, t, u, v, w, phi, scheme)
        assert rel_err(t[g.interior], -c * div) < 1e-12


def test_weno_z5_1d_advection_order(orc):
    # uniform u = 1, sin profile along x: tend -> -dϕ/dx at 5th order
    errs = []
    for N in (16, 32):
        g = orc.Grid((N, 4, 4), (2 * np.pi, 1.0, 1.0), halo=3)
        u = np.ones(g.storage, order="F")
        v = g.zeros(); w = g.zeros()
        x = (np.arange(g.storage[0]) - 3 + 0.5) * g.dx
        # cell averages of sin
        phi = np.asfortranarray(np.broadcast_to(((np.cos(x - g.dx / 2) - np.cos(x + g.dx / 2)) / g.dx)[:, None, None],
                                                g.storage).copy())
        t = g.zeros()
        orc.advect_scalar(g, t, u, v, w, phi, "weno_z5")
        xi = x[3:3 + N]
        exact = -(np.sin(xi + g.dx / 2) - np.sin(xi - g.dx / 2)) / g.dx  # -(flux difference of the exact face values)
        errs.append(np.abs(t[3:3 + N, 3, 3] - exact).max())
    assert math.log2(errs[0] / errs[1]) > 4.5


def test_centered_generic_equals_centered_operator(orc):
    # Centered2 through the generic reconstruction path equals the centred
    # scalar operator bit for bit (the WENO composition's zero-order baseline)
    g = orc.Grid((8, 7, 6), (1.0, 1.0, 1.0), halo=3)
    u, v, w, phi = (rand_storage(g.storage, s) for s in (11, 12, 13, 14))
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ), (phi, orc.CC)):
        orc.fill_halos(g, a, loc, ("periodic",) * 3)
    t1 = g.zeros()
    orc.advect_scalar(g, t1, u, v, w, phi, "centered")
    # a numpy restatement of _adv_centered_at (src/FiniteVolume.jl:1682-1693)
    I = slice(3, -3)
    Ip, Im = slice(4, -2), slice(2, -4)
    d = phi
    r = ((u[Ip, I, I] * 0.5 * (d[I, I, I] + d[Ip, I, I]) - u[I, I, I] * 0.5 * (d[Im, I, I] + d[I, I, I])) / g.dx
         + (v[I, Ip, I] * 0.5 * (d[I, I, I] + d[I, Ip, I]) - v[I, I, I] * 0.5 * (d[I, Im, I] + d[I, I, I])) / g.dy
         + (w[I, I, Ip] * 0.5 * (d[I, I, I] + d[I, I, Ip]) - w[I, I, I] * 0.5 * (d[I, I, Im] + d[I, I, I])) / g.dz)
    assert rel_err(t1[I, I, I], -r) < 1e-14


# ---------------------------------------------------------------------------
# canonical reductions: the 8-lane transcript (test/backends.jl:30-43)
# ---------------------------------------------------------------------------
def reference_plane_sum(A, k):
    l = [0.0] * 8
    n = A.shape[0]
    nb = n // 8
    for j in range(A.shape[1]):
        for b in range(nb):
            for r in range(8):
                l[r] += A[8 * b + r, j, k]
        for i in range(8 * nb, n):
            l[0] += A[i, j, k]
    return ((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + (l[6] + l[7]))


def test_canonical_sum_transcript(orc):
    A = rand_storage((13, 5, 4), 3)
    s = 0.0
    for k in range(4):
        s += reference_plane_sum(A, k)
    assert orc.interior_sum(A) == s
    assert abs(orc.interior_sum(A) - math.fsum(A.ravel())) < 1e-14 * np.abs(A).sum()


# ---------------------------------------------------------------------------
# Poisson solvers: exact inverses of the discrete operator
# ---------------------------------------------------------------------------
def _residual(orc, g, phi, src, bcs):
    H = g.H
    P = g.zeros()
    P[g.interior] = phi
    orc.fill_halos(g, P, orc.CC, bcs)
    lap = orc.laplacian(g, P, orc.CC)[g.interior]
    return lap - (src - src.mean()), phi


def test_fft_solver_inverts_laplacian(orc):
    g = orc.Grid((16, 8, 32), (1.0, 2.0, 3.0), halo=1)
    src = np.asfortranarray(np.random.default_rng(2).normal(size=(16, 8, 32)))
    phi = orc.fft_solver(g).solve(src)
    r, _ = _residual(orc, g, phi, src, ("periodic",) * 3)
    assert np.abs(r).max() < 1e-10 * np.abs(src).max()
    assert abs(phi.mean()) < 1e-12 * np.abs(phi).max()


@pytest.mark.parametrize("stretched", [False, True])
def test_column_solver_inverts_laplacian(orc, stretched):
    from aeolus_hip.fields import faces_geometric  # face generator only (host numpy)
    Nz = 12
    zf = faces_geometric(Nz, 0.0, 100.0, 1.08) if stretched else None
    g = orc.Grid((8, 16, Nz), (40.0, 80.0, 100.0), halo=1, zf=zf)
    src = np.asfortranarray(np.random.default_rng(4).normal(size=(8, 16, Nz)))
    dzc, _ = g.interior_dz()
    src -= (src * dzc).sum() / (dzc.sum() * 8 * 16)  # compatible: zero volume-weighted mean
    phi = orc.stretched_column_solver(g).solve(src)
    P = g.zeros()
    P[g.interior] = phi
    orc.fill_halos(g, P, orc.CC, ("periodic", "periodic", "noflux"))
    lap = orc.laplacian(g, P, orc.CC)[g.interior]
    assert np.abs(lap - src).max() < 1e-10 * np.abs(src).max()
    # the gauge: zero volume-weighted mean
    assert abs((phi * dzc).sum()) < 1e-11 * np.abs(phi).max() * phi.size * dzc.max()


def test_column_solver_matches_dense(orc):
    # brute force on a tiny grid: the dense 7-point operator (periodic x-y,
    # Neumann z, stretched), solved in the zero-volume-mean gauge
    from aeolus_hip.fields import faces_geometric
    Nx, Ny, Nz = 4, 4, 5
    zf = faces_geometric(Nz, 0.0, 10.0, 1.3)
    g = orc.Grid((Nx, Ny, Nz), (4.0, 6.0, 10.0), halo=1, zf=zf)
    dzc, dzf = g.interior_dz()
    n = Nx * Ny * Nz
    A = np.zeros((n, n))
    lin = lambda i, j, k: (i % Nx) + Nx * ((j % Ny) + Ny * k)
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                r = lin(i, j, k)
                for di, d in ((1, g.dx), (-1, g.dx)):
                    A[r, lin(i + di, j, k)] += 1 / d ** 2; A[r, r] -= 1 / d ** 2
                for dj, d in ((1, g.dy), (-1, g.dy)):
                    A[r, lin(i, j + dj, k)] += 1 / d ** 2; A[r, r] -= 1 / d ** 2
                if k < Nz - 1:
                    c = 1 / (dzc[k] * dzf[k + 1]); A[r, lin(i, j, k + 1)] += c; A[r, r] -= c
                if k > 0:
                    c = 1 / (dzc[k] * dzf[k]); A[r, lin(i, j, k - 1)] += c; A[r, r] -= c
    src = np.random.default_rng(9).normal(size=(Nx, Ny, Nz))
    wgt = np.broadcast_to(dzc[None, None, :], src.shape)
    src -= (src * wgt).sum() / wgt.sum()
    b = src.reshape(-1, order="F")
    # gauge row: Σ dzc φ = 0
    Ag = np.vstack([A, wgt.reshape(-1, order="F")])
    x = np.linalg.lstsq(Ag, np.append(b, 0.0), rcond=None)[0].reshape((Nx, Ny, Nz), order="F")
    phi = orc.stretched_column_solver(g).solve(np.asfortranarray(src))
    assert rel_err(phi, x) < 1e-11


def test_column_incompatible_source_residual_is_minus_mean(orc):
    # test/pressure_column.jl:44-56 (uniform grid): L φ − s = −s̄ uniformly
    g = orc.Grid((8, 8, 6), (1.0, 1.0, 1.0), halo=1)
    raw = np.asfortranarray(np.random.default_rng(5).normal(size=(8, 8, 6)) + 0.3)
    phi = orc.stretched_column_solver(g).solve(raw)
    P = g.zeros(); P[g.interior] = phi
    orc.fill_halos(g, P, orc.CC, ("periodic", "periodic", "noflux"))
    r = orc.laplacian(g, P, orc.CC)[g.interior] - raw
    assert np.abs(r + raw.mean()).max() < 1e-10


# ---------------------------------------------------------------------------
# the step
# ---------------------------------------------------------------------------
def _tgv(orc, N, dt, steps, nu, U=1.0, scheme="centered"):
    from aeolus_hip.workloads import tgv_state
    g = orc.Grid((N, N, N), (2 * np.pi,) * 3, halo=3)
    u, v, w, th = tgv_state(g, U)
    p = g.zeros()
    st = orc.Stepper(g, scheme=scheme, nu=nu)
    for _ in range(steps):
        st.rk3(u, v, w, th, p, dt)
    return g, u, v, w


def _ke(g, u, v, w):
    I = g.interior
    return 0.5 * ((u[I] ** 2).sum() + (v[I] ** 2).sum() + (w[I] ** 2).sum()) * g.dx * g.dy * g.dz


def test_step_kills_divergence(orc):
    # test/ka_step.jl:118-123: max |∇·u| < 1e-12 after a step from a random start
    g = orc.Grid((16, 16, 16), (2 * np.pi,) * 3, halo=3)
    u, v, w = (rand_storage(g.storage, s) for s in (404, 405, 406))
    th = np.full(g.storage, 300.0, order="F"); p = g.zeros()
    st = orc.Stepper(g, scheme="weno_z5", nu=1e-2)
    st.euler(u, v, w, th, p, 0.05)
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(g, a, loc, ("periodic",) * 3)
    assert orc.max_divergence(g, u, v, w) < 1e-12


def test_rk3_third_order(orc):
    # test/ka_step.jl:124-139: error ratio in [6.5, 9.5] when dt halves
    N, nu = 16, 0.05
    Kref = _ke(*_tgv(orc, N, 1 / 80, 80, nu))
    e1 = abs(_ke(*_tgv(orc, N, 0.1, 10, nu)) - Kref)
    e2 = abs(_ke(*_tgv(orc, N, 0.05, 20, nu)) - Kref)
    assert 6.5 < e1 / e2 < 9.5


def test_tgv_decays_at_discrete_rate(orc):
    # test/ka_step.jl:140-146: K(t) = K0 exp(-2ν k̃² t), k̃² = 3 (2/Δ sin(Δ/2))²
    N, nu, U = 16, 0.05, 1e-3
    g0 = _tgv(orc, N, 0.05, 0, nu, U)
    K0 = _ke(*g0)
    g1 = _tgv(orc, N, 0.05, 20, nu, U)
    D = 2 * np.pi / N
    k2 = 3 * (2 / D * np.sin(D / 2)) ** 2
    assert abs(_ke(*g1) / K0 / np.exp(-2 * nu * k2) - 1) < 1e-4


def test_dcbl_budget_and_walls(orc):
    # flux form: Σ θ Δz_c changes only by the surface flux F_bot · Δt per column;
    # w stays zero on the walls; the projected field is solenoidal
    from aeolus_hip.workloads import dcbl_faces, dcbl_state
    Nz = 16
    zf, Lz = dcbl_faces(Nz)
    g = orc.Grid((16, 16, Nz), (320.0, 320.0, Lz), halo=3, zf=zf)
    u, v, w, th = dcbl_state(g)
    rng = np.random.default_rng(1)
    I = g.interior
    u[I] += 0.5 * rng.uniform(-1, 1, size=u[I].shape)
    v[I] += 0.5 * rng.uniform(-1, 1, size=v[I].shape)
    p = g.zeros()
    st = orc.Stepper(g, scheme="weno_z5", theta0=300.0, flux_bot=0.1, bcs=("periodic", "periodic", "wall"),
                     theta_bcs=("periodic", "periodic", "noflux"))
    dzc, _ = g.interior_dz()
    dt = 0.5
    st.euler(u, v, w, th, p, dt)  # project the random start: every RK3 stage is then solenoidal
    m0 = (th[I] * dzc[None, None, :]).sum()
    st.rk3(u, v, w, th, p, dt)
    m1 = (th[I] * dzc[None, None, :]).sum()
    assert abs((m1 - m0) - 0.1 * dt * 16 * 16) < 1e-9 * abs(m0)
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(g, a, loc, ("periodic", "periodic", "wall"))
    assert np.all(w[:, :, 3] == 0)
    assert orc.max_divergence(g, u, v, w) < 1e-12 * max(1.0, np.abs(u).max() / g.dx)


# ---------------------------------------------------------------------------
# multigrid (CPU oracle): the reference's test/multigrid.jl checks
# ---------------------------------------------------------------------------
def _orc_mg(orc, N, spacing, bcs, coeff=None, rhs_scale=None, smoother=0):
    L = orc.lib()
    cs = [None] * 3 if coeff is None else [orc.ptr(np.asfortranarray(c)) for c in coeff]
    rs = None if rhs_scale is None else orc.ptr(np.ascontiguousarray(rhs_scale))
    return L.orc_mg_create(*N, *spacing, *[orc.BCS[b] for b in bcs], *cs, rs, 2, 2, 4, smoother, 0.8)


def test_mg_manufactured_and_levels(orc):
    # test/multigrid.jl:10-31: ≥ 3 levels, recovery of a manufactured solution
    import ctypes
    N = (32, 32, 16)
    sp = (1.0 / 32, 1.0 / 32, 0.5 / 16)
    h = _orc_mg(orc, N, sp, ("periodic", "periodic", "neumann"))
    L = orc.lib()
    assert L.orc_mg_levels(h) >= 3
    x = (np.arange(32) + 0.5) / 32
    z = (np.arange(16) + 0.5) * sp[2]
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    phi = np.asfortranarray(np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) * np.cos(np.pi * Z / 0.5))
    phi -= phi.mean()
    b = np.zeros(N, order="F")
    L.orc_mg_apply(h, orc.ptr(b), orc.ptr(phi))
    out = np.zeros(N, order="F")
    rel = ctypes.c_double()
    cyc = L.orc_mg_solve(h, orc.ptr(out), orc.ptr(b), 1e-12, 60, ctypes.byref(rel))
    assert rel.value < 1e-12 and cyc < 30
    assert np.abs(out - phi).max() < 1e-9
    L.orc_mg_destroy(h)


def test_mg_w_cycle_and_fmg(orc):
    # cycle = :W and fmg = true (src/Multigrid.jl:559-570, 706-760): every
    # variant converges to the direct column solver's solution; a W-cycle
    # contracts more per cycle than a V-cycle (fewer cycles); the FMG start
    # leaves a first-cycle residual far below a zero start's
    import ctypes
    N = (32, 32, 16)
    g = orc.Grid(N, (32.0, 32.0, 16.0), halo=1)
    src = np.asfortranarray(np.random.default_rng(5).normal(size=N))
    src -= src.mean()
    x_dir = orc.stretched_column_solver(g).solve(src)
    L = orc.lib()
    res = {}
    for w in (0, 1):
        for f in (0, 1):
            h = _orc_mg(orc, N, (1.0, 1.0, 1.0), ("periodic", "periodic", "neumann"))
            L.orc_mg_set_cycle(h, w, f)
            out = np.zeros(N, order="F")
            rel = ctypes.c_double()
            cyc = L.orc_mg_solve(h, orc.ptr(out), orc.ptr(src), 1e-11, 100, ctypes.byref(rel))
            one = np.zeros(N, order="F")
            rel1 = ctypes.c_double()
            L.orc_mg_solve(h, orc.ptr(one), orc.ptr(src), 1e-30, 1, ctypes.byref(rel1))
            L.orc_mg_destroy(h)
            assert rel.value < 1e-11
            assert np.abs(out - x_dir).max() < 1e-8 * np.abs(x_dir).max()
            res[(w, f)] = (cyc, rel1.value)
    assert res[(1, 0)][0] < res[(0, 0)][0]
    assert res[(0, 1)][1] < 0.1 * res[(0, 0)][1] and res[(1, 1)][1] < 0.1 * res[(1, 0)][1]
    assert res[(0, 1)][0] < res[(0, 0)][0]


def test_mgpcg_matches_column_solver(orc):
    # MG-CG and the direct column solver invert the same operator (test/multigrid.jl:23-31)
    import ctypes
    N = (16, 16, 8)
    g = orc.Grid(N, (16.0, 16.0, 8.0), halo=1)
    src = np.asfortranarray(np.random.default_rng(1).normal(size=N))
    src -= src.mean()
    x_dir = orc.stretched_column_solver(g).solve(src)
    h = _orc_mg(orc, N, (1.0, 1.0, 1.0), ("periodic", "periodic", "neumann"))
    L = orc.lib()
    x = np.zeros(N, order="F")
    it = ctypes.c_int(); res = ctypes.c_double(); bn = ctypes.c_double()
    st = L.orc_mgpcg_solve(h, orc.ptr(x), orc.ptr(src), 1e-12, 100, ctypes.byref(it), ctypes.byref(res),
                           ctypes.byref(bn))
    L.orc_mg_destroy(h)
    assert st == 0
    assert rel_err(x, x_dir) < 1e-9


def test_mgpcg_rhs_scale_invariance(orc):
    # test/multigrid.jl:75-83: scaling the source scales the solution
    import ctypes
    N = (16, 16, 8)
    src = np.asfortranarray(np.random.default_rng(2).normal(size=N))
    sols = []
    for s in (1.0, 1e3):
        h = _orc_mg(orc, N, (1.0, 1.0, 1.0), ("periodic", "periodic", "neumann"))
        x = np.zeros(N, order="F")
        it = ctypes.c_int(); res = ctypes.c_double(); bn = ctypes.c_double()
        orc.lib().orc_mgpcg_solve(h, orc.ptr(x), orc.ptr(np.asfortranarray(s * src)), 1e-12, 100, ctypes.byref(it),
                                  ctypes.byref(res), ctypes.byref(bn))
        orc.lib().orc_mg_destroy(h)
        sols.append(x / s)
    assert rel_err(sols[1], sols[0]) < 1e-8


def test_cutcell_geometry_matches_oracle(orc):
    # the product's numpy builder (host setup) against the oracle's C restatement, bit for bit
    from aeolus_hip.cutcell import heightmap_geometry
    N, H = (24, 20, 16), 2
    g = orc.Grid(N, (48.0, 40.0, 32.0), halo=H)
    rng = np.random.default_rng(3)
    h = np.zeros((N[0], N[1]))
    for _ in range(5):
        i0, j0 = rng.integers(0, 18), rng.integers(0, 14)
        h[i0:i0 + rng.integers(2, 6), j0:j0 + rng.integers(2, 6)] = rng.uniform(3.0, 25.0)
    vol, ve, ax, ay, az = heightmap_geometry(N, (2.0, 2.0, 2.0), H, h, nsub=4)
    sx, sy, sz = g.storage
    ii = np.clip(np.arange(1, sx + 1) - H, 1, N[0]) - 1
    jj = np.clip(np.arange(1, sy + 1) - H, 1, N[1]) - 1
    hfull = np.asfortranarray(h[np.ix_(ii, jj)])
    out = [np.zeros(g.storage, order="F") for _ in range(5)]
    orc.lib().orc_cutcell_heightmap(ctypes_byref(g), orc.ptr(hfull), 4, 1e-3, *[orc.ptr(a) for a in out])
    for a, b in zip((vol, ve, ax, ay, az), out):
        assert np.array_equal(a, b)
    # a flat roof at 5.0 m cuts the cells [4, 6) m to half volume
    assert vol[H + 0, H + 0, H + 2] in (0.0, 0.5, 1.0)


def _small_cell_case(orc, seed=3, vol_min=0.3):
    # roofs at heights that leave a quarter of a cell fluid: with vol_min 0.3
    # those cells are small (the default vol_min = 1e-3 never is with a
    # heightmap: a cell's fraction is a multiple of 1/nsub)
    from aeolus_hip.cutcell import heightmap_geometry
    N, H = (24, 20, 16), 2
    g = orc.Grid(N, (48.0, 40.0, 32.0), halo=H)
    rng = np.random.default_rng(seed)
    h = np.zeros((N[0], N[1]))
    for _ in range(6):
        i0, j0 = rng.integers(0, 18), rng.integers(0, 14)
        h[i0:i0 + rng.integers(2, 6), j0:j0 + rng.integers(2, 6)] = 2.0 * rng.integers(1, 10) + 1.4
    vol, ve, ax, ay, az = heightmap_geometry(N, (2.0, 2.0, 2.0), H, h, nsub=4, vol_min=vol_min)
    return g, N, H, vol, ve, ax, ay, az


def test_small_cell_merges_match_oracle(orc):
    # the product's merge list (host setup) against the oracle's restatement
    # of _finalize_geometry's partner rule (src/CutCell.jl:466-488)
    from aeolus_hip.cutcell import small_cell_merges
    g, N, H, vol, ve, ax, ay, az = _small_cell_case(orc)
    pairs = small_cell_merges(vol, ve, ax, ay, az, N, H, 0.3)
    t = orc.cutcell_merge_targets(g, vol, ve, ax, ay, az, 0.3)
    inner = np.zeros(g.storage, bool)
    inner[g.interior] = True
    flat = t.ravel(order="F")
    want = [(c, flat[c] - 1) for c in np.flatnonzero((inner & (t > 0)).ravel(order="F"))]
    # the oracle lists cells in Fortran (i fastest) order: the loop order
    assert len(pairs) > 10
    assert [tuple(p) for p in pairs] == want


def test_redistribution_conserves_and_merges(orc):
    # cutcell_divergence!(…; redistribute = true): Σ vol_eff·div is kept (the
    # cluster mean is vol_eff-weighted), each pair ends equal, and cells in no
    # cluster keep the strict D
    g, N, H, vol, ve, ax, ay, az = _small_cell_case(orc)
    t = orc.cutcell_merge_targets(g, vol, ve, ax, ay, az, 0.3)
    rng = np.random.default_rng(9)
    u, v, w = (np.asfortranarray(rng.standard_normal(g.storage)) for _ in range(3))
    d0 = orc.cutcell_divergence(g, u, v, w, ax, ay, az, ve, t, redistribute=False)
    d1 = orc.cutcell_divergence(g, u, v, w, ax, ay, az, ve, t, redistribute=True)
    sx, sy, sz = g.storage
    I = g.interior
    # partners inside the interior: the conserved sum is over the interior
    tgt = t[I][t[I] > 0] - 1
    ti, tj, tk = tgt % sx, (tgt // sx) % sy, tgt // (sx * sy)
    assert np.all((ti >= H) & (ti < H + N[0]) & (tj >= H) & (tj < H + N[1]) & (tk >= H) & (tk < H + N[2]))
    s0, s1 = (ve[I] * d0[I]).sum(), (ve[I] * d1[I]).sum()
    assert abs(s1 - s0) < 1e-12 * np.abs(ve[I] * d0[I]).sum()
    flat1 = d1.ravel(order="F")
    last = {}
    for c in np.flatnonzero(t.ravel(order="F") * np.asarray(np.pad(np.ones(N, bool), H), dtype=np.int64).ravel(order="F")):
        last[c] = flat1[c]
    touched = set(last) | set((t.ravel(order="F")[list(last)] - 1).tolist())
    untouched = np.ones(sx * sy * sz, bool)
    untouched[list(touched)] = False
    m = untouched & np.asarray(np.pad(np.ones(N, bool), H)).ravel(order="F")
    assert np.array_equal(flat1[m], d0.ravel(order="F")[m])
    assert len(touched) > 10


def ctypes_byref(g):
    import ctypes
    return ctypes.byref(g.c)


# ---------------------------------------------------------------------------
# golden vectors: WENO face values in exact rational arithmetic from the
# polynomial / smoothness-indicator definitions (tests/golden/make_golden.py)
# ---------------------------------------------------------------------------
def test_weno_faces_match_exact_rational_golden(orc):
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "weno_faces.json")
    cases = json.load(open(path))["cases"]
    L = orc.lib()
    worst = 0.0
    for c in cases:
        v = [float.fromhex(x) for x in c["v7"]]
        scale = max(abs(x) for x in v)
        got = {
            "weno_z5_p1": L.orc_weno_z5_left(*v[1:6], 1),
            "weno_z5_p2": L.orc_weno_z5_left(*v[1:6], 2),
            "weno_z7_p2": L.orc_weno_z7_left(*v, 2),
            "weno_js5": L.orc_weno_js5_left(*v[1:6], 1e-6),
        }
        for k, g in got.items():
            e = abs(g - float.fromhex(c[k])) / scale
            worst = max(worst, e)
            assert e < 1e-13, (k, v, g, float.fromhex(c[k]))
    print(f"worst WENO face error vs exact rational: {worst:.2e} (relative to the stencil's max)")


@pytest.mark.parametrize("bcx,bcy", [("neumann", "periodic"), ("periodic", "neumann"), ("neumann", "neumann")])
def test_column_solver_neumann_lateral_inverts_laplacian(orc, bcx, bcy):
    # the REDFT10 path: the 7-point Laplacian with mirror (Neumann) lateral
    # neighbours and the stretched z operator, applied in numpy to the
    # solution, returns the source (minus its measure-weighted mean)
    from aeolus_hip.fields import faces_geometric
    N = (16, 12, 10)
    zf = faces_geometric(N[2], 0.0, 50.0, 1.08)
    g = orc.Grid(N, (32.0, 36.0, 50.0), halo=1, zf=zf)
    dzc, dzf = g.interior_dz()
    src = np.random.default_rng(5).normal(size=N)
    src -= (src * dzc).sum() / (dzc.sum() * N[0] * N[1])
    phi = orc.stretched_column_solver(g, bcx, bcy).solve(np.asfortranarray(src))

    def nb(a, axis, bc, shift):
        if bc == "periodic":
            return np.roll(a, shift, axis=axis)
        pad = [(0, 0)] * 3
        pad[axis] = (1, 1)
        e = np.pad(a, pad, mode="edge")
        sl = [slice(None)] * 3
        sl[axis] = slice(1 - shift, 1 - shift + a.shape[axis])
        return e[tuple(sl)]
    lap = (nb(phi, 0, bcx, 1) - 2 * phi + nb(phi, 0, bcx, -1)) / g.dx ** 2
    lap += (nb(phi, 1, bcy, 1) - 2 * phi + nb(phi, 1, bcy, -1)) / g.dy ** 2
    up_ = np.zeros_like(phi)
    dn = np.zeros_like(phi)
    up_[:, :, :-1] = (phi[:, :, 1:] - phi[:, :, :-1]) / dzf[1:-1]
    dn[:, :, 1:] = (phi[:, :, 1:] - phi[:, :, :-1]) / dzf[1:-1]
    lap += (up_ - dn) / dzc
    assert np.abs(lap - src).max() < 1e-10 * np.abs(src).max()
