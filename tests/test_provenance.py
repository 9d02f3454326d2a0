"""provenance() (src/io/run_folder.jl:149-260) on the CPU: the git state of
this checkout as git reports it, the library's SHA-256 of its bytes and its
build string, the AEOLUS_* environment, the command."""
import hashlib
import os
import subprocess

import pytest


def test_provenance_record(monkeypatch):
    from aeolus_hip import _lib as L
    from aeolus_hip.provenance import SOURCE_DIR, git_state, provenance
    monkeypatch.setenv("AEOLUS_TEST_FLAG", "1")
    rec = provenance()
    assert set(rec) >= {"git", "library", "host", "python", "packages", "command", "env"}
    assert rec["env"]["AEOLUS_TEST_FLAG"] == "1"
    assert rec["command"]["cwd"] == os.getcwd()
    g = rec["git"]
    r = subprocess.run(["git", "-C", SOURCE_DIR, "rev-parse", "HEAD"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no git checkout")
    assert g["available"] and g["commit"] == r.stdout.strip()
    st = subprocess.run(["git", "-C", SOURCE_DIR, "status", "--porcelain", "--untracked-files=all"],
                        capture_output=True, text=True).stdout
    assert g["dirty"] == bool(st.strip())
    # excluding every changed path leaves the tree clean
    paths = sorted({l[3:].strip('"').split("/")[0] for l in st.split("\n") if l.strip()})
    assert not git_state(SOURCE_DIR, exclude=paths)["dirty"]
    lib = rec["library"]
    if os.path.exists(L.LIB_PATH):
        assert lib["sha256"] == hashlib.sha256(open(L.LIB_PATH, "rb").read()).hexdigest()
        assert "gfx950" in lib["build"] and "HIP" in lib["build"]


def test_git_state_outside_a_checkout(tmp_path):
    from aeolus_hip.provenance import git_state
    g = git_state(str(tmp_path))
    assert g["available"] is False and g["commit"] == "" and not g["dirty"]
