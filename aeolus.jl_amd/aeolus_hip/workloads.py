"""Seeded synthetic inputs of the benchmark and test workloads (host numpy).

tgv:  Taylor-Green vortex, triply periodic [0, 2π)^3, u = sin x cos y cos z,
      v = -cos x sin y cos z, w = 0, Re = 1600 (ν = 1/1600), θ uniform.
dcbl: dry convective boundary layer, periodic x-y, walls in z, stretched z
      (Δz 10 m at the surface growing 2 % per level), θ = 300 K up to 1 km
      then +3 K/km, seeded uniform ±0.1 K noise in the lowest 10 levels,
      surface heat flux 0.1 K m/s, Δx = Δy = 20 m.
"""
import numpy as np

from .fields import Grid, faces_geometric

TGV_SEED = 1600
DCBL_SEED = 20260


def tgv_grid(N=64, halo=3):
    L = 2 * np.pi
    return Grid((N, N, N), (L, L, L), halo=halo)


def tgv_state(grid, U=1.0):
    """u, v, w, θ on the storage (halo points evaluated analytically, then
    overwritten by the first halo fill)."""
    H = grid.H
    sx, sy, sz = grid.storage
    i = np.arange(sx) - H
    j = np.arange(sy) - H
    k = np.arange(sz) - H
    xf = i * grid.dx                  # x of FaceX
    xc = (i + 0.5) * grid.dx          # x of cell centres
    yf = j * grid.dy
    yc = (j + 0.5) * grid.dy
    zc = (k + 0.5) * grid.dz
    X, Y, Z = np.meshgrid(xf, yc, zc, indexing="ij")
    u = U * np.sin(X) * np.cos(Y) * np.cos(Z)
    X, Y, Z = np.meshgrid(xc, yf, zc, indexing="ij")
    v = -U * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros(grid.storage)
    th = np.full(grid.storage, 300.0)
    return [np.asfortranarray(a) for a in (u, v, w, th)]


def dcbl_faces(Nz=128, dz1=10.0, r=1.02):
    """Interior faces of the stretched column: geometric with Δz₁ = dz1, ratio r
    (_faces_geometric with Lz = dz1 (r^Nz - 1)/(r - 1))."""
    Lz = dz1 * (r ** Nz - 1) / (r - 1)
    return faces_geometric(Nz, 0.0, Lz, r), Lz


def dcbl_grid(N=(256, 256, 128), dx=20.0, halo=3):
    zf, Lz = dcbl_faces(N[2])
    return Grid(N, (N[0] * dx, N[1] * dx, Lz), halo=halo, zf=zf)


def dcbl_state(grid, seed=DCBL_SEED):
    H = grid.H
    zf = grid.zf
    zc = 0.5 * (zf[:-1] + zf[1:])
    prof = np.where(zc <= 1000.0, 300.0, 300.0 + 0.003 * (zc - 1000.0))
    th = np.zeros(grid.storage)
    th[:, :, H:H + grid.Nz] = prof[None, None, :]
    rng = np.random.default_rng(seed)
    noise = rng.uniform(-0.1, 0.1, size=(grid.Nx, grid.Ny, 10))
    th[H:H + grid.Nx, H:H + grid.Ny, H:H + 10] += noise
    u = np.zeros(grid.storage)
    v = np.zeros(grid.storage)
    w = np.zeros(grid.storage)
    return [np.asfortranarray(a) for a in (u, v, w, th)]


DCBL_PHYS = dict(nu=0.0, theta0=300.0, g_accel=9.81, flux_bot=0.1, flux_top=0.0,
                 bcs=("periodic", "periodic", "wall"), theta_bcs=("periodic", "periodic", "noflux"))
TGV_PHYS = dict(nu=1.0 / 1600, theta0=0.0, bcs=("periodic",) * 3)
