"""A restart from step N continues the run bit for bit on the GPU
(test/io_checkpoint.jl's contract, src/io/checkpoint.jl:473-480), for the
Boussinesq core with Coriolis and a sponge and for the sound-proof core."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _boussinesq(hip):
    from aeolus_hip import workloads as W
    zf, Lz = W.dcbl_faces(24)
    g = hip.Grid((32, 32, 24), (640.0, 640.0, Lz), halo=3, zf=zf)
    sim = hip.BoussinesqSimulation(g, scheme="weno_z5", f_coriolis=1e-3, sponge_alpha_max=0.02,
                                   sponge_z_start=0.6 * Lz, sponge_target=dict(u=2.0), **W.DCBL_PHYS)
    state = W.dcbl_state(g)
    rng = np.random.default_rng(7)
    for a in state[:2]:
        a[g.interior] += 0.5 * rng.uniform(-1, 1, size=a[g.interior].shape)
    sim.set(*state)
    return sim, lambda: hip.BoussinesqSimulation(g, scheme="weno_z5", f_coriolis=1e-3, sponge_alpha_max=0.02,
                                                 sponge_z_start=0.6 * Lz, sponge_target=dict(u=2.0),
                                                 **W.DCBL_PHYS), 0.5


def _anelastic(hip):
    from aeolus_hip.anelastic import AnelasticSimulation, isothermal_reference
    N = (32, 16, 20)
    g = hip.Grid(N, (N[0] * 200.0, N[1] * 200.0, N[2] * 100.0), halo=1)
    ref = isothermal_reference(N[2], 100.0, 1)
    sim = AnelasticSimulation(g, ref, nu=1.0)
    rng = np.random.default_rng(3)
    st = g.storage
    th = np.zeros(st, order="F")
    th[:, :, 1:1 + N[2]] = ref.theta[None, None, :]
    u, v, w = (np.zeros(st, order="F") for _ in range(3))
    th[g.interior] += 0.3 * rng.standard_normal(th[g.interior].shape)
    u[g.interior] += rng.standard_normal(u[g.interior].shape)
    sim.set(u, v, w, th)
    return sim, lambda: AnelasticSimulation(g, ref, nu=1.0), 3.0


@pytest.mark.parametrize("core", ["boussinesq", "anelastic"])
def test_restart_continues_bit_for_bit(hip, tmp_path, core):
    sim, make, dt = (_boussinesq if core == "boussinesq" else _anelastic)(hip)
    for _ in range(2):
        sim.step_rk3(dt)
    path = hip.checkpoint(sim, str(tmp_path / "ck.npz"))
    for _ in range(2):
        sim.step_rk3(dt)
    cont = sim.state()
    other = make()
    hip.restart(other, path)
    assert other.t == 2 * dt
    for _ in range(2):
        other.step_rk3(dt)
    got = other.state()
    for n in ("u", "v", "w", "theta", "p"):
        assert np.array_equal(got[n], cont[n]), n
    assert other.t == sim.t


def test_provenance_names_the_device_and_library(hip, tmp_path):
    from aeolus_hip.provenance import device_record, provenance
    d = device_record(0)
    print(f"device: {d}", flush=True)
    # the marketing name may be empty when the driver reports none; the
    # architecture, compute units and memory identify the part
    assert d["arch"].startswith("gfx9") and d["compute_units"] > 0 and d["memory_bytes"] > 0
    rec = provenance(device=0)
    assert rec["device"] == d and len(rec["library"]["sha256"]) == 64
    # a checkpoint carries the provenance of the run that wrote it
    g = hip.Grid((8, 8, 8), (1.0, 1.0, 1.0), halo=3)
    sim = hip.BoussinesqSimulation(g)
    path = str(tmp_path / "c.npz")
    hip.checkpoint(sim, path)
    meta, _ = hip.read_checkpoint(path)
    assert meta["provenance"]["library"]["sha256"] == rec["library"]["sha256"]
