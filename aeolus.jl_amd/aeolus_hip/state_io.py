"""Full-state checkpoint and restart of a device-resident run.

Follows the reference's versioned checkpoint (src/io/checkpoint.jl:346-560):
the prognostic fields with their halos (u, v, w, p, θ) in the storage layout,
the clock, the mesh record, the run's scalar parameters, the base state of a
sound-proof core, and provenance. `restart` refuses an unversioned or newer
file, a mesh, domain or core mismatch, and differing run parameters unless
`allow_mismatch` (then it warns), as restart! does (:430-560). A restart
from step N then continues the run bit for bit: every solve of the library
starts from zero and the step keeps no other state between calls.

The container is a NumPy .npz (no pickles: the metadata is one JSON string)
written through `path + ".tmp"` and a rename, so an interrupted write never
leaves a truncated checkpoint (write_checkpoint, :383-392). The reference
serialises a Julia Dict, which is not readable outside Julia; the field set,
layout and checks are the same.
"""
import json
import os
import time
import warnings

import numpy as np

from .provenance import provenance

CHECKPOINT_FORMAT = "Aeolus.checkpoint"
CHECKPOINT_VERSION = 1
_FIELDS = ("u", "v", "w", "p", "theta")


def _mesh_record(grid):
    """_mesh_record (src/io/checkpoint.jl:291-294): sizes, halo, extent,
    spacings; a stretched grid records its cell thicknesses."""
    rec = dict(Nx=grid.Nx, Ny=grid.Ny, Nz=grid.Nz, H=grid.H, Lx=grid.Lx, Ly=grid.Ly, Lz=grid.Lz,
               dx=grid.dx, dy=grid.dy)
    if getattr(grid, "zf", None) is None:
        rec["dz"] = grid.dz
    else:
        rec["dz"] = np.diff(np.asarray(grid.zf, dtype=np.float64)).tolist()
    return rec


def _jsonable(x):
    if isinstance(x, np.ndarray):
        return x.tolist()
    if isinstance(x, (tuple, list)):
        return [_jsonable(a) for a in x]
    if isinstance(x, dict):
        return {k: _jsonable(v) for k, v in x.items()}
    if isinstance(x, np.generic):
        return x.item()
    return x


def checkpoint_payload(sim, extra=None):
    """The in-memory checkpoint (checkpoint_payload, :303-345): fields are
    host copies, nothing aliases the run."""
    fields = {n: np.asfortranarray(getattr(sim, n).download()) for n in _FIELDS}
    meta = dict(format=CHECKPOINT_FORMAT, version=CHECKPOINT_VERSION, created=time.time(),
                sim_type=type(sim).__name__, mesh=_mesh_record(sim.grid), t=float(getattr(sim, "t", 0.0)),
                parameters=_jsonable(sim.parameters()), reference=_jsonable(checkpoint_reference(sim)),
                precision=int(sim.precision), extra=_jsonable(extra or {}), provenance=provenance())
    return meta, fields


def checkpoint_reference(sim):
    """checkpoint_reference (:283-289): nothing for the Boussinesq core, the
    base-state profiles of a sound-proof core."""
    ref = getattr(sim, "ref", None)
    if ref is None:
        return None
    return dict(rho=ref.rho, rho_f=ref.rho_f, theta=ref.theta, theta_f=ref.theta_f)


def write_checkpoint(path, meta, fields):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        np.savez(f, meta=np.array(json.dumps(meta)), **{"field_" + k: v for k, v in fields.items()})
    os.replace(tmp, path)
    return path


def checkpoint(sim, path, extra=None):
    """checkpoint!(sim, path) (:346-350)."""
    meta, fields = checkpoint_payload(sim, extra)
    return write_checkpoint(path, meta, fields)


def read_checkpoint(path):
    """read_checkpoint (:394-425): the (meta, fields) of a versioned file;
    ValueError for an unversioned file or a newer format."""
    with np.load(path, allow_pickle=False) as z:
        if "meta" not in z.files:
            raise ValueError(f"restart: {path} is not a versioned checkpoint (no metadata)")
        meta = json.loads(str(z["meta"]))
        if not isinstance(meta, dict) or meta.get("format") != CHECKPOINT_FORMAT:
            raise ValueError(f"restart: {path} is an unversioned snapshot; a run cannot be restarted from it")
        v = meta.get("version")
        if not isinstance(v, int) or v > CHECKPOINT_VERSION:
            raise ValueError(f"restart: {path} has checkpoint format version {v}; this build reads up to "
                             f"version {CHECKPOINT_VERSION}")
        fields = {k[len("field_"):]: z[k] for k in z.files if k.startswith("field_")}
    return meta, fields


def _equal(a, b):
    return json.dumps(a, sort_keys=True) == json.dumps(b, sort_keys=True)


def restart(sim, path, allow_mismatch=False):
    """restart!(sim, path; allow_mismatch) (:473-560): sim must be built as
    the checkpointed one; copies the fields with their halos and the clock."""
    meta, fields = read_checkpoint(path)
    m, fm = _mesh_record(sim.grid), meta["mesh"]
    sizes = ("Nx", "Ny", "Nz", "H")
    if tuple(fm[k] for k in sizes) != tuple(m[k] for k in sizes):
        raise ValueError(f"restart: mesh mismatch: file {tuple(fm[k] for k in sizes)} vs. sim "
                         f"{tuple(m[k] for k in sizes)} (Nx, Ny, Nz, halo)")
    if (fm["Lx"], fm["Ly"], fm["Lz"]) != (m["Lx"], m["Ly"], m["Lz"]):
        raise ValueError(f"restart: domain mismatch: file {(fm['Lx'], fm['Ly'], fm['Lz'])} vs. sim "
                         f"{(m['Lx'], m['Ly'], m['Lz'])}")
    if not _equal(fm["dz"], m["dz"]):
        raise ValueError("restart: mesh mismatch: the vertical spacings differ")
    if meta["sim_type"] != type(sim).__name__:
        raise ValueError(f"restart: checkpoint of a {meta['sim_type']}, restoring into a {type(sim).__name__}")
    if int(meta["precision"]) != int(sim.precision):
        raise ValueError("restart: the checkpoint's state precision differs from the simulation's")
    p, fp = _jsonable(sim.parameters()), dict(meta["parameters"])
    diff = sorted(k for k in fp if k in p and not _equal(p[k], fp[k]))
    if not _equal(_jsonable(checkpoint_reference(sim)), meta.get("reference")):
        diff.append("reference")
    if diff:
        listing = ", ".join(f"{k} = {p.get(k, 'base state')} (checkpoint: {fp.get(k, 'base state')})"
                            for k in diff)
        if not allow_mismatch:
            raise ValueError(f"restart: run parameters differ from the checkpointed run in {path}: {listing}. "
                             "Build the simulation with the checkpoint's parameters, or pass "
                             "allow_mismatch=True to restart anyway")
        warnings.warn(f"restart: restarting with allow_mismatch=True; run parameters differ: {listing}")
    for n in _FIELDS:
        dev = getattr(sim, n)
        a = fields[n]
        if a.shape != tuple(sim.grid.storage):
            raise ValueError(f"restart: field {n} has shape {a.shape}, the simulation's storage is "
                             f"{tuple(sim.grid.storage)}")
        dev.upload(np.asfortranarray(a))
    sim.t = float(meta["t"])
    return meta
