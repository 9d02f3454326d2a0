"""What produced a run (provenance, src/io/run_folder.jl:149-260): the git
state of the source tree, the library that ran (its path, SHA-256 and build
string), the device, the host, the Python packages, the command and the
AEOLUS_* environment, as one JSON-ready dict. bench.py puts it in its result
line and checkpoints carry it in their metadata.

The reference fingerprints its package Manifest with FNV-1a; here the
fingerprint of the code that ran is the SHA-256 of the shared library, and
the packages are the versions of the Python modules the host code imports.
"""
import ctypes
import hashlib
import os
import platform
import socket
import subprocess
import sys

from . import _lib as L

SOURCE_DIR = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def _git(repo, *args):
    try:
        r = subprocess.run(["git", "-C", repo, *args], capture_output=True, text=True, timeout=10)
    except (OSError, subprocess.SubprocessError):
        return False, ""
    return r.returncode == 0, r.stdout


def git_state(repo=SOURCE_DIR, exclude=()):
    """git_state(repo; exclude) (src/io/run_folder.jl:156-181): available,
    toplevel, commit, branch, dirty and the `git status --porcelain` lines,
    without paths under `exclude` (relative to the top level)."""
    ok, top = _git(repo, "rev-parse", "--show-toplevel")
    if not ok:
        return dict(available=False, toplevel="", commit="", branch="", dirty=False, changes=[])
    top = top.strip()
    _, commit = _git(repo, "rev-parse", "HEAD")
    _, branch = _git(repo, "rev-parse", "--abbrev-ref", "HEAD")
    _, status = _git(repo, "status", "--porcelain", "--untracked-files=all")
    changes = [l for l in status.split("\n") if l.strip()]
    ex = [e.rstrip("/") for e in exclude]
    changes = [l for l in changes if not any(l[3:].strip('"') == e or l[3:].strip('"').startswith(e + "/") for e in ex)]
    return dict(available=True, toplevel=top, commit=commit.strip(), branch=branch.strip(), dirty=bool(changes),
                changes=changes)


def library_record():
    """The shared library that runs the kernels: path, SHA-256, build string."""
    path = L.LIB_PATH
    rec = dict(path=path, sha256="", build="")
    if os.path.exists(path):
        h = hashlib.sha256()
        with open(path, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 20), b""):
                h.update(chunk)
        rec["sha256"] = h.hexdigest()
        rec["build"] = L.load().aeo_build_info().decode()
    return rec


def device_record(device=0):
    """The GPU: name, gfx architecture, compute units, HBM bytes (initialises
    the library on `device`)."""
    L.init(device)
    name, arch = ctypes.create_string_buffer(256), ctypes.create_string_buffer(64)
    cus, mem = ctypes.c_int(), ctypes.c_size_t()
    L.check(L.load().aeo_device_info(device, name, 256, arch, 64, ctypes.byref(cus), ctypes.byref(mem)))
    return dict(index=device, name=name.value.decode(), arch=arch.value.decode(), compute_units=cus.value,
                memory_bytes=mem.value)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def provenance(repo=SOURCE_DIR, exclude=(), device=None):
    """provenance(; repo, exclude) (src/io/run_folder.jl:215-260) as a dict:
    git, library, device (when `device` is given), host, python, packages,
    command and env."""
    pk = {}
    for m in ("numpy", "torch", "scipy"):
        mod = sys.modules.get(m)
        if mod is not None:
            pk[m] = getattr(mod, "__version__", "")
    rec = dict(
        git=git_state(repo, exclude),
        library=library_record(),
        host=dict(hostname=socket.gethostname(), os=platform.system() + " " + platform.release(),
                  arch=platform.machine(), cpu=_cpu_model(), cpu_threads=os.cpu_count(),
                  omp_threads=os.environ.get("OMP_NUM_THREADS", "")),
        python=dict(version=platform.python_version(), executable=sys.executable),
        packages=pk,
        command=dict(program=sys.argv[0] if sys.argv else "", args=list(sys.argv[1:]), cwd=os.getcwd()),
        env={k: v for k, v in os.environ.items() if k.startswith("AEOLUS_")},
    )
    if device is not None:
        rec["device"] = device_record(device)
    return rec
