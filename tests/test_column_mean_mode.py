"""The mean mode of the column solve (_mean_mode!, src/PressureSolvers.jl:
780-805) against its closed form. A horizontally uniform source excites only
the (kx, ky) = (0, 0) column, so the whole solve is the mean-mode path: the
source minus its Δz_c-weighted mean, integrated up as a flux and gauged to
Σ φ Δz_c = 0. The reference is a dense solve of the 1-D operator
(φ_{k+1} − φ_k)/(Δz_c Δz_f,k+1) − (φ_k − φ_{k−1})/(Δz_c Δz_f,k) with
Neumann ends, by least squares, shifted to the same gauge. The GPU's
mean-mode column (one wave, scans) is held to the same closed form."""
import numpy as np
import pytest


def _dense_mean_mode(dzc, dzf, f):
    Nz = len(dzc)
    A = np.zeros((Nz, Nz))
    for k in range(Nz):
        lo = 1 / (dzc[k] * dzf[k]) if k > 0 else 0.0
        up = 1 / (dzc[k] * dzf[k + 1]) if k < Nz - 1 else 0.0
        if k > 0:
            A[k, k - 1] = lo
        if k < Nz - 1:
            A[k, k + 1] = up
        A[k, k] = -(lo + up)
    rhs = f - (f * dzc).sum() / dzc.sum()
    phi = np.linalg.lstsq(A, rhs, rcond=None)[0]
    phi -= (phi * dzc).sum() / dzc.sum()
    # least squares on the stretched operator is good to ~cond·eps, not to
    # round-off: 1e-10 still separates a right answer from a wrong gauge or a
    # missing mean subtraction, which are O(1) errors
    assert np.abs(A @ phi - rhs).max() < 1e-10 * np.abs(rhs).max()
    return phi


def _case(Nz=40, seed=3):
    from aeolus_hip.workloads import dcbl_faces
    zf, Lz = dcbl_faces(Nz)
    f = np.random.default_rng(seed).standard_normal(Nz) * 1e-2
    return zf, Lz, f


def test_oracle_mean_mode_closed_form(orc):
    zf, Lz, f = _case()
    og = orc.Grid((12, 10, len(f)), (240.0, 200.0, Lz), halo=3, zf=zf)
    dzc, dzf = og.interior_dz()
    want = _dense_mean_mode(dzc, dzf, f)
    s = orc.stretched_column_solver(og)
    src = np.asfortranarray(np.broadcast_to(f, (12, 10, len(f))).copy())
    phi = s.solve(src)
    assert np.abs(phi - want[None, None, :]).max() <= 1e-10 * np.abs(want).max()


@pytest.mark.gpu
@pytest.mark.parametrize("Nz", [40, 150])
def test_gpu_mean_mode_closed_form(hip, Nz):
    # Nz = 150: the wave walks three chunks of 64 levels, the last one partial
    zf, Lz, f = _case(Nz, seed=Nz)
    g = hip.Grid((32, 16, Nz), (640.0, 320.0, Lz), halo=3, zf=zf)
    dzc, dzf = g.interior_dz()
    want = _dense_mean_mode(dzc, dzf, f)
    s = hip.solver_for(g)
    phi = hip.solve_interior(s, np.broadcast_to(f, (32, 16, Nz)).copy())
    assert np.abs(phi - want[None, None, :]).max() <= 1e-10 * np.abs(want).max()
