"""advect_positive! on the GPU against the oracle (src/AdvectionSchemes.jl:
758-1160): every reconstruction, periodic and wall boundaries, the limiter's
halo from the BCs or the one-layer wrap, a stretched grid (centred), FP32
error stated; the reference's narrow-Gaussian case (test/advection.jl:92-119)
run on the device; positivity and conservation at 256×256×128."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _case(hip, orc, N, H, bcs, seed, zf=None, ext=None):
    ext = ext or tuple(float(n) for n in N)
    g = hip.Grid(N, ext, halo=H, zf=zf)
    og = orc.Grid(N, ext, halo=H, zf=zf)
    rng = np.random.default_rng(seed)
    u, v, w = (og.zeros() for _ in range(3))
    for a in (u, v, w):
        a[...] = rng.standard_normal(a.shape)
    for a, loc in ((u, orc.FX), (v, orc.FY), (w, orc.FZ)):
        orc.fill_halos(og, a, loc, bcs)
    phi = og.zeros()
    phi[...] = rng.uniform(0, 1, phi.shape) ** 4  # sharp, positive: the limiter acts
    tbcs = tuple("noflux" if b == "wall" else b for b in bcs)
    orc.fill_halos(og, phi, orc.CC, tbcs)
    return g, og, (u, v, w, phi), tbcs


def _up(hip, a, prec=0):
    from aeolus_hip.fields import DeviceArray
    return DeviceArray.from_numpy(a.astype(np.float32) if prec else a, prec)


@pytest.mark.parametrize("scheme,p", [("weno_z5", 2), ("weno_z5", 1), ("weno_z7", 2), ("weno5", 2),
                                      ("weno_js5", 2), ("centered", 2)])
@pytest.mark.parametrize("walls", [False, True])
@pytest.mark.parametrize("use_bcs", [False, True])
def test_advect_positive_parity(hip, orc, scheme, p, walls, use_bcs):
    H = 4 if scheme == "weno_z7" else 3
    bcs = ("periodic", "periodic", "wall") if walls else ("periodic",) * 3
    g, og, (u, v, w, phi), tbcs = _case(hip, orc, (40, 24, 20), H, bcs, 5 + walls)
    dt = 0.15
    rng = np.random.default_rng(2)
    tend0 = og.zeros()
    tend0[...] = rng.standard_normal(tend0.shape)
    ref = orc.advect_positive(og, tend0.copy(order="F"), u, v, w, phi, dt, scheme=scheme, p=p,
                              bcs=tbcs if use_bcs else None)
    dt_ = _up(hip, tend0)
    hip.advect_positive(g, dt_, *(_up(hip, a) for a in (u, v, w, phi)), dt, scheme=scheme, p=p,
                        bcs=tbcs if use_bcs else None)
    got = dt_.download()
    I = og.interior
    assert rel_err(got[I] - tend0[I], ref[I] - tend0[I]) <= TOL64
    assert np.array_equal(got[0], tend0[0])  # the halo is not touched
    # the limiter had work to do: unlimited WENO/centred would undershoot here
    raw = orc.advect_scalar(og, og.zeros(), u, v, w, phi, "weno_js5" if scheme == "weno5" else scheme)
    assert np.abs(raw[I] - (ref[I] - tend0[I])).max() > 1e-3 * np.abs(raw[I]).max()


def test_advect_positive_stretched_centered(hip, orc):
    from aeolus_hip import workloads as W
    zf, Lz = W.dcbl_faces(20)
    bcs = ("periodic", "periodic", "wall")
    g, og, (u, v, w, phi), tbcs = _case(hip, orc, (24, 16, 20), 3, bcs, 12, zf=zf, ext=(480.0, 320.0, Lz))
    for use_bcs in (False, True):
        ref = orc.advect_positive(og, og.zeros(), u, v, w, phi, 2.0, scheme="centered",
                                  bcs=tbcs if use_bcs else None, accumulate=False)
        t = _up(hip, og.zeros())
        hip.advect_positive(g, t, *(_up(hip, a) for a in (u, v, w, phi)), 2.0, scheme="centered",
                            bcs=tbcs if use_bcs else None, accumulate=False)
        I = og.interior
        assert rel_err(t.download()[I], ref[I]) <= TOL64
    with pytest.raises(hip.AeolusError, match="centered only"):
        hip.advect_positive(g, t, *(_up(hip, a) for a in (u, v, w, phi)), 2.0, scheme="weno_z5")


def test_workspace_reuse_and_errors(hip, orc):
    bcs = ("periodic",) * 3
    g, og, arrs, _ = _case(hip, orc, (32, 16, 12), 3, bcs, 3)
    d = [_up(hip, a) for a in arrs]
    ws = hip.FluxLimiterWorkspace(g)
    outs = []
    for wsp in (ws, ws, None):
        t = _up(hip, og.zeros())
        hip.advect_positive(g, t, *d, 0.1, workspace=wsp, accumulate=False)
        outs.append(t.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    with pytest.raises(ValueError):
        hip.advect_positive(g, t, *d, 0.1, scheme="upwind")
    g2 = hip.Grid((32, 16, 12), (32.0, 16.0, 12.0), halo=3)
    with pytest.raises(ValueError):
        hip.advect_positive(g2, t, *d, 0.1, workspace=ws)


def test_narrow_gaussian_on_the_device(hip, orc):
    # test/advection.jl:92-119: 20 forward-Euler steps at CFL 0.95
    N, h = 16, 1.0 / 16
    g = hip.Grid((N,) * 3, (1.0,) * 3, halo=3)
    og = orc.Grid((N,) * 3, (1.0,) * 3, halo=3)
    bcs = ("periodic",) * 3
    u, v, w = (og.zeros() for _ in range(3))
    u[...], v[...], w[...] = 1.0, 0.6, 0.3
    dt = 0.95 / (1.0 / h + 0.6 / h + 0.3 / h)
    s = 1.2 * h
    x = (np.arange(N) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    I = og.interior
    phi = og.zeros()
    phi[I] = np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) / (2 * s * s))
    m0 = phi[I].sum()
    du, dv, dw = (_up(hip, a) for a in (u, v, w))
    dphi, t = _up(hip, phi), _up(hip, og.zeros())
    ws = hip.FluxLimiterWorkspace(g)
    ref = phi.copy(order="F")
    for _ in range(20):
        hip.fill_halos(g, dphi, 0, bcs)
        hip.advect_positive(g, t, du, dv, dw, dphi, dt, workspace=ws, accumulate=False)
        a, tt = dphi.download(), t.download()
        a[I] += dt * tt[I]
        dphi.upload(a)
        orc.fill_halos(og, ref, orc.CC, bcs)
        r = orc.advect_positive(og, og.zeros(), u, v, w, ref, dt, accumulate=False)
        ref[I] += dt * r[I]
    got = dphi.download()[I]
    assert got.min() >= -1e-14 * got.max()
    assert abs((got.sum() - m0) / m0) < 1e-12
    assert rel_err(got, ref[I]) <= 1e-11


def test_advect_positive_fp32_error_stated(hip, orc):
    bcs = ("periodic", "periodic", "wall")
    g, og, (u, v, w, phi), tbcs = _case(hip, orc, (40, 24, 20), 3, bcs, 8)
    ref = orc.advect_positive(og, og.zeros(), u, v, w, phi, 0.15, bcs=tbcs, accumulate=False)
    t = _up(hip, og.zeros(), 1)
    hip.advect_positive(g, t, *(_up(hip, a, 1) for a in (u, v, w, phi)), 0.15, bcs=tbcs, accumulate=False)
    I = og.interior
    err = rel_err(t.download()[I].astype(np.float64), ref[I])
    print(f"advect_positive FP32 vs FP64 oracle: {err:.2e}")
    assert err < 1e-4


def test_full_size_positivity_and_conservation(hip):
    # 256x256x128, periodic: one limited Euler update of a sharp positive field
    # at first-order CFL 0.45 stays non-negative and keeps the total
    from aeolus_hip.fields import DeviceArray
    N = (256, 256, 128)
    g = hip.Grid(N, (256.0, 256.0, 128.0), halo=3)
    rng = np.random.default_rng(77)
    st = g.storage
    I = g.interior
    bcs = ("periodic",) * 3
    dev = []
    for loc in (1, 2, 3):
        a = np.asfortranarray(rng.standard_normal(st))
        d = DeviceArray.from_numpy(a)
        hip.fill_halos(g, d, loc, bcs)
        dev.append(d)
    phi = np.asfortranarray(rng.uniform(0, 1, st) ** 6)
    dphi = DeviceArray.from_numpy(phi)
    hip.fill_halos(g, dphi, 0, bcs)
    phi = dphi.download()
    umax = max(np.abs(d.download()).max() for d in dev)
    dt = 0.45 / (3 * umax) / 2
    t = DeviceArray(st)
    hip.advect_positive(g, t, *dev, dphi, dt, workspace=hip.FluxLimiterWorkspace(g), accumulate=False)
    tt = t.download()[I]
    new = phi[I] + dt * tt
    assert new.min() >= -1e-15 * phi[I].max()
    assert abs(tt.sum()) <= 1e-11 * np.abs(tt).sum()
