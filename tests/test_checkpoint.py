"""Checkpoint container and restart checks on the host (no GPU): the field
set round-trips with its halos and layout, and restart refuses what the
reference's restart! refuses (src/io/checkpoint.jl:394-560)."""
import json
import os
import warnings

import numpy as np
import pytest

from aeolus_hip import state_io as C


class _HostArray:
    """download/upload over a host array: the DeviceArray surface the
    checkpoint code uses."""

    def __init__(self, shape):
        self.a = np.zeros(shape, order="F")

    def download(self):
        return self.a.copy(order="F")

    def upload(self, a):
        self.a[...] = a


class _Grid:
    def __init__(self, N=(6, 5, 4), H=3, extent=(60.0, 50.0, 40.0), zf=None):
        self.Nx, self.Ny, self.Nz = N
        self.H = H
        self.Lx, self.Ly, self.Lz = extent
        self.dx, self.dy, self.dz = extent[0] / N[0], extent[1] / N[1], extent[2] / N[2]
        self.zf = zf
        self.storage = tuple(n + 2 * H for n in N)


class BoussinesqSimulation:
    def __init__(self, grid, nu=0.1, precision=0):
        self.grid, self.precision, self.t = grid, precision, 0.0
        self.ref = None
        self._p = dict(nu=nu, theta0=300.0, bcs=["periodic", "periodic", "wall"])
        for n in ("u", "v", "w", "p", "theta"):
            setattr(self, n, _HostArray(grid.storage))

    def parameters(self):
        return dict(self._p)


def _filled(seed=0, **kw):
    sim = BoussinesqSimulation(_Grid(), **kw)
    rng = np.random.default_rng(seed)
    for n in ("u", "v", "w", "p", "theta"):
        getattr(sim, n).a[...] = rng.standard_normal(sim.grid.storage)
    sim.t = 12.5
    return sim


def test_round_trip_bit_exact(tmp_path):
    a = _filled()
    path = str(tmp_path / "run" / "ck.npz")
    C.checkpoint(a, path, extra=dict(note="x"))
    assert os.path.exists(path) and not os.path.exists(path + ".tmp")
    b = BoussinesqSimulation(_Grid())
    meta = C.restart(b, path)
    for n in ("u", "v", "w", "p", "theta"):
        assert np.array_equal(getattr(a, n).a, getattr(b, n).a), n
        assert getattr(b, n).a.flags.f_contiguous
    assert b.t == 12.5 and meta["extra"] == dict(note="x") and meta["version"] == C.CHECKPOINT_VERSION


def test_refuses_unversioned_and_newer(tmp_path):
    sim = _filled()
    p1 = str(tmp_path / "snap.npz")
    np.savez(p1, field_u=np.zeros(3))
    with pytest.raises(ValueError, match="not a versioned checkpoint"):
        C.restart(sim, p1)
    meta, fields = C.checkpoint_payload(sim)
    meta["version"] = C.CHECKPOINT_VERSION + 1
    p2 = C.write_checkpoint(str(tmp_path / "new.npz"), meta, fields)
    with pytest.raises(ValueError, match="format version"):
        C.restart(sim, p2)
    meta["version"], meta["format"] = 1, "other"
    p3 = C.write_checkpoint(str(tmp_path / "other.npz"), meta, fields)
    with pytest.raises(ValueError, match="unversioned"):
        C.restart(sim, p3)


def test_refuses_mesh_domain_core_mismatch(tmp_path):
    path = C.checkpoint(_filled(), str(tmp_path / "ck.npz"))
    with pytest.raises(ValueError, match="mesh mismatch"):
        C.restart(BoussinesqSimulation(_Grid(N=(6, 5, 5))), path)
    with pytest.raises(ValueError, match="domain mismatch"):
        C.restart(BoussinesqSimulation(_Grid(extent=(60.0, 50.0, 41.0))), path)
    with pytest.raises(ValueError, match="mesh mismatch"):
        C.restart(BoussinesqSimulation(_Grid(zf=np.array([0.0, 5.0, 15.0, 25.0, 40.0]))), path)
    with pytest.raises(ValueError, match="precision"):
        C.restart(BoussinesqSimulation(_Grid(), precision=1), path)

    class AnelasticSimulation(BoussinesqSimulation):
        pass
    with pytest.raises(ValueError, match="restoring into a AnelasticSimulation"):
        C.restart(AnelasticSimulation(_Grid()), path)


def test_parameter_mismatch_refused_or_warned(tmp_path):
    a = _filled()
    path = C.checkpoint(a, str(tmp_path / "ck.npz"))
    b = BoussinesqSimulation(_Grid(), nu=0.2)
    with pytest.raises(ValueError, match=r"nu = 0.2 \(checkpoint: 0.1\)"):
        C.restart(b, path)
    assert np.abs(b.u.a).max() == 0  # nothing restored
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        C.restart(b, path, allow_mismatch=True)
    assert any("allow_mismatch" in str(x.message) for x in w)
    assert np.array_equal(a.theta.a, b.theta.a)


def test_metadata_is_json_without_pickles(tmp_path):
    path = C.checkpoint(_filled(), str(tmp_path / "ck.npz"))
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        assert set(f[len("field_"):] for f in z.files if f.startswith("field_")) == {"u", "v", "w", "p", "theta"}
    assert meta["format"] == "Aeolus.checkpoint" and meta["mesh"]["H"] == 3
