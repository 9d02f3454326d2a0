"""Cut-cell geometry of flat-roofed buildings and its pressure operator (host setup).

cut_cell_geometry_from_heightmap (src/CutCell.jl:372-396) through
cut_cell_geometry's sub-cell sampling (:213-312) and _finalize_geometry's
consistency sweeps and small-cell floor (:398-480), vectorised in numpy:
the level set is φ(x, y, z) = z − h(column), φ > 0 fluid. The small-cell
merge partners are built (small_cell_merges) for the divergence's
redistribution; the surface normals are not (no operator here reads them).

The pressure operator of the projection is D·G (aperture power 2,
cutcell_laplacian!, :735-781); in the multigrid's face-coefficient form the
symmetric operator vol_eff·L has c = a² on every face (0 on blocked faces)
and the source vol_eff·(D u) is the aperture-weighted face-flux balance.
"""
import ctypes

import numpy as np


def _col(x, d, H, n):
    """The storage column containing x (1-based, clamped): idx in src/CutCell.jl:385."""
    return np.clip(np.floor(x / d).astype(np.int64) + H + 1, 1, n)


def heightmap_geometry(N, spacing, H, height, nsub=4, vol_min=1e-3):
    """(vol, vol_eff, ax, ay, az) on the storage (Fortran arrays), for interior
    size N = (Nx, Ny, Nz), spacing (dx, dy, dz), halo H, origin 0 and `height`
    of shape (Nx, Ny) (edge-replicated into the halo) or the storage (sx, sy)."""
    Nx, Ny, Nz = N
    dx, dy, dz = spacing
    sx, sy, sz = Nx + 2 * H, Ny + 2 * H, Nz + 2 * H
    height = np.asarray(height, dtype=np.float64)
    if height.shape == (sx, sy):
        hfull = height
    elif height.shape == (Nx, Ny):
        ii = np.clip(np.arange(1, sx + 1) - H, 1, Nx) - 1
        jj = np.clip(np.arange(1, sy + 1) - H, 1, Ny) - 1
        hfull = height[np.ix_(ii, jj)]
    else:
        raise ValueError(f"height must be {Nx}x{Ny} or {sx}x{sy}, got {height.shape}")
    invn = 1.0 / nsub
    hx, hy, hz = dx * invn, dy * invn, dz * invn
    volnorm, areanorm = 1.0 / nsub ** 3, 1.0 / nsub ** 2
    i = np.arange(1, sx + 1)
    j = np.arange(1, sy + 1)
    k = np.arange(1, sz + 1)
    xb = (i - H - 1) * dx          # _xw
    yb = (j - H - 1) * dy
    zb = (k - H - 1) * dz
    sub = np.arange(1, nsub + 1) - 0.5
    xs = xb[:, None] + sub[None, :] * hx       # (sx, nsub)
    ys = yb[:, None] + sub[None, :] * hy
    cxs = _col(xs, dx, H, sx)                  # columns of the sub-samples
    cys = _col(ys, dy, H, sy)
    cxf = _col(xb, dx, H, sx)                  # columns at the faces
    cyf = _col(yb, dy, H, sy)
    # levels whose cells or bottom faces can be cut: below the tallest roof
    # (every sample above it is fluid; the sweeps below cannot reach past
    # the first all-fluid level, whose cells keep open faces)
    hmax = float(hfull.max())
    kc = int(np.searchsorted(zb, hmax, side="right")) + 1
    kc = min(max(kc, 1), sz)
    zs = zb[:kc, None] + sub[None, :] * hz     # (kc, nsub)
    zbk = zb[:kc]
    vol = np.ones((sx, sy, sz))
    ax = np.ones((sx, sy, sz))
    ay = np.ones((sx, sy, sz))
    az = np.ones((sx, sy, sz))
    vk = np.zeros((sx, sy, kc))
    axk = np.zeros((sx, sy, kc))
    ayk = np.zeros((sx, sy, kc))
    azk = np.zeros((sx, sy, kc))
    # counts of fluid sub-samples: a loop over the horizontal sub-samples, vectorised over cells
    for a in range(nsub):
        for b in range(nsub):
            hab = hfull[np.ix_(cxs[:, a] - 1, cys[:, b] - 1)]            # (sx, sy)
            for c in range(nsub):
                vk += zs[None, None, :, c] - hab[:, :, None] > 0
            azk += zbk[None, None, :] - hab[:, :, None] > 0
        hxa = hfull[np.ix_(cxf - 1, cys[:, a] - 1)]                        # x-face columns
        hya = hfull[np.ix_(cxs[:, a] - 1, cyf - 1)]                        # y-face columns
        for c in range(nsub):
            axk += zs[None, None, :, c] - hxa[:, :, None] > 0
            ayk += zs[None, None, :, c] - hya[:, :, None] > 0
    vol[:, :, :kc] = vk * volnorm
    ax[:, :, :kc] = axk * areanorm
    ay[:, :, :kc] = ayk * areanorm
    az[:, :, :kc] = azk * areanorm
    del vk, axk, ayk, azk
    # _finalize_geometry (on the levels that can hold solid cells, plus one)
    ke = min(kc + 1, sz)
    V, AX, AY, AZ = vol[:, :, :ke], ax[:, :, :ke], ay[:, :, :ke], az[:, :, :ke]
    solid = ~(V > 0)
    ax[0, :, :] = 0.0
    ay[:, 0, :] = 0.0
    az[:, :, 0] = 0.0
    for _ in range(16):
        AX[1:, :, :][solid[:-1, :, :] | solid[1:, :, :]] = 0.0
        AY[:, 1:, :][solid[:, :-1, :] | solid[:, 1:, :]] = 0.0
        AZ[:, :, 1:][solid[:, :, :-1] | solid[:, :, 1:]] = 0.0
        opn = (AX + AY) + AZ
        opn[:-1, :, :] += AX[1:, :, :]
        opn[:, :-1, :] += AY[:, 1:, :]
        opn[:, :, :-1] += AZ[:, :, 1:]
        if ke < sz:
            opn[:, :, -1] += az[:, :, ke]
        newly = ~solid & ~(opn > 0)
        if not newly.any():
            break
        solid |= newly
        V[newly] = 0.0
    vol_eff = np.where(vol < vol_min, vol_min, vol)
    vol_eff[:, :, :ke][solid] = 0.0
    F = lambda a: np.asfortranarray(a)
    return F(vol), F(vol_eff), F(ax), F(ay), F(az)


def small_cell_merges(vol, vol_eff, ax, ay, az, N, H, vol_min=1e-3):
    """The small cells (not solid, vol < vol_min) of the interior and their
    merge partners, as an int32 array of (cell, target) 0-based storage
    linear indices (Fortran order), in the interior loop order of
    _redistribute_small_cells! (src/CutCell.jl:628-648). The partner is the
    neighbour of largest vol through an open face, faces taken −x, +x, −y,
    +y, −z, +z, the first strictly largest kept (src/CutCell.jl:466-488);
    small cells without one are left out."""
    sx, sy, sz = vol.shape
    small = (vol_eff > 0) & (vol < vol_min)
    Nx, Ny, Nz = N
    inner = np.zeros_like(small)
    inner[H:H + Nx, H:H + Ny, H:H + Nz] = True
    cells = np.argwhere(small & inner)
    # the loop order of `for k, j, i`: k slowest
    cells = cells[np.lexsort((cells[:, 0], cells[:, 1], cells[:, 2]))]
    pairs = []
    for i, j, k in cells:
        cand = ((ax[i, j, k] if i > 0 else 0.0, i - 1, j, k), (ax[i + 1, j, k] if i < sx - 1 else 0.0, i + 1, j, k),
                (ay[i, j, k] if j > 0 else 0.0, i, j - 1, k), (ay[i, j + 1, k] if j < sy - 1 else 0.0, i, j + 1, k),
                (az[i, j, k] if k > 0 else 0.0, i, j, k - 1), (az[i, j, k + 1] if k < sz - 1 else 0.0, i, j, k + 1))
        best, bestv = -1, 0.0
        for ap, i2, j2, k2 in cand:
            if not ap > 0 or not (0 <= i2 < sx and 0 <= j2 < sy and 0 <= k2 < sz):
                continue
            v2 = vol[i2, j2, k2]
            if v2 > bestv:
                bestv, best = v2, i2 + sx * (j2 + sy * k2)
        if best >= 0:
            pairs.append((i + sx * (j + sy * k), best))
    return np.asarray(pairs, dtype=np.int32).reshape(-1, 2)


def mg_coefficients(ax, ay, az, N, H):
    """Face coefficients a² of the D·G operator on the interior faces:
    cx (Nx+1, Ny, Nz), cy (Nx, Ny+1, Nz), cz (Nx, Ny, Nz+1)."""
    Nx, Ny, Nz = N
    I = slice(H, H + Nx)
    J = slice(H, H + Ny)
    K = slice(H, H + Nz)
    cx = ax[H:H + Nx + 1, J, K] ** 2
    cy = ay[I, H:H + Ny + 1, K] ** 2
    cz = az[I, J, H:H + Nz + 1] ** 2
    return np.asfortranarray(cx), np.asfortranarray(cy), np.asfortranarray(cz)


def cutcell_source(u, v, w, ax, ay, az, vol_eff, N, H, spacing):
    """vol_eff · (D u): the aperture-weighted face-flux balance of the face
    velocities (cutcell_divergence!, src/CutCell.jl:602-626, times the cell
    volume fraction), zero in solid cells, over the interior."""
    Nx, Ny, Nz = N
    dx, dy, dz = spacing
    I = slice(H, H + Nx); J = slice(H, H + Ny); K = slice(H, H + Nz)
    Ip = slice(H + 1, H + Nx + 1); Jp = slice(H + 1, H + Ny + 1); Kp = slice(H + 1, H + Nz + 1)
    fx = (ax[Ip, J, K] * u[Ip, J, K] - ax[I, J, K] * u[I, J, K]) / dx
    fy = (ay[I, Jp, K] * v[I, Jp, K] - ay[I, J, K] * v[I, J, K]) / dy
    fz = (az[I, J, Kp] * w[I, J, Kp] - az[I, J, K] * w[I, J, K]) / dz
    b = (fx + fy) + fz
    b[~(vol_eff[I, J, K] > 0)] = 0.0
    return np.asfortranarray(b)


# ---------------------------------------------------------------------------
# the operators on the GPU (cutcell.hip through the C ABI)
# ---------------------------------------------------------------------------
class CutCellGeometry:
    """CutCellGeometry (src/CutCell.jl:177-195) on the device: apertures ax,
    ay, az and the effective volume fractions vol_eff (FP64 storage arrays),
    built on the host by heightmap_geometry."""

    def __init__(self, grid, height, nsub=4, vol_min=1e-3):
        from .fields import DeviceArray
        N = (grid.Nx, grid.Ny, grid.Nz)
        vol, ve, ax, ay, az = heightmap_geometry(N, (grid.dx, grid.dy, grid.dz), grid.H, height, nsub, vol_min)
        self.grid = grid
        self.host = dict(vol=vol, vol_eff=ve, ax=ax, ay=ay, az=az)
        self.ax, self.ay, self.az, self.vol_eff = (DeviceArray.from_numpy(a) for a in (ax, ay, az, ve))
        self.merges = small_cell_merges(vol, ve, ax, ay, az, N, grid.H, vol_min)
        self._merges_dev = None
        if len(self.merges):
            from . import _lib as L
            p = ctypes.c_void_p()
            nbytes = self.merges.nbytes
            L.check(L.init().aeo_malloc(ctypes.byref(p), nbytes))
            L.check(L.init().aeo_memcpy_h2d(p, self.merges.ctypes.data_as(ctypes.c_void_p), nbytes))
            self._merges_dev = p.value

    def __del__(self):
        try:
            from . import _lib as L
            if self._merges_dev and L._lib is not None and L._initialised:
                L._lib.aeo_free(ctypes.c_void_p(self._merges_dev))
        except Exception:
            pass

    def coefficients(self):
        """Face coefficients a² of vol_eff·D·G for the multigrid solvers."""
        h = self.host
        return mg_coefficients(h["ax"], h["ay"], h["az"], (self.grid.Nx, self.grid.Ny, self.grid.Nz), self.grid.H)

    def fluid_mask(self):
        H, g = self.grid.H, self.grid
        return np.asfortranarray((self.host["vol_eff"][H:H + g.Nx, H:H + g.Ny, H:H + g.Nz] > 0).astype(np.float64))

    def _geo(self):
        from . import _lib as L
        return L.dptr(self.ax), L.dptr(self.ay), L.dptr(self.az)


def divergence(div, u, v, w, geom, redistribute=True):
    """cutcell_divergence!(div, u, v, w, geom; redistribute) (src/CutCell.jl:602-648):
    with redistribute, each small cell and its merge partner take the
    cluster's vol_eff-weighted mean divergence."""
    from . import _lib as L
    L.check(L.init().aeo_cutcell_divergence(geom.grid.handle, div.precision, L.dptr(div), L.dptr(u), L.dptr(v),
                                            L.dptr(w), *geom._geo(), L.dptr(geom.vol_eff)))
    if redistribute and geom._merges_dev:
        L.check(L.init().aeo_cutcell_redistribute(geom.grid.handle, div.precision, L.dptr(div), L.dptr(geom.vol_eff),
                                                  ctypes.c_void_p(geom._merges_dev), len(geom.merges)))
    return div


def gradient(gx, gy, gz, p, geom):
    """cutcell_gradient_x!/_y!/_z! (src/CutCell.jl:666-700); p FP64 with its halo filled."""
    from . import _lib as L
    L.check(L.init().aeo_cutcell_gradient(geom.grid.handle, gx.precision, L.dptr(gx), L.dptr(gy), L.dptr(gz),
                                          L.dptr(p), *geom._geo()))
    return gx, gy, gz


def laplacian(out, f, geom, aperture_power=2):
    """cutcell_laplacian!(out, f, geom; aperture_power) (src/CutCell.jl:735-781)."""
    from . import _lib as L
    L.check(L.init().aeo_cutcell_laplacian(geom.grid.handle, out.precision, L.dptr(out), L.dptr(f), *geom._geo(),
                                           L.dptr(geom.vol_eff), aperture_power))
    return out


def project(u, v, w, geom, solver, bcs=("periodic", "periodic", "neumann")):
    """The cut-cell projection: solve vol_eff·D·G φ = vol_eff·D u* with
    `solver` (an MGPreconditionedCG on geom.coefficients()), then u ← u* − G φ.
    Returns φ (FP64 interior-sized DeviceArray)."""
    from . import _lib as L
    from .fields import DeviceArray
    from .pressure import solve_source_buffer
    g = geom.grid
    L.check(L.init().aeo_cutcell_project_source(g.handle, u.precision, solver.handle, L.dptr(u), L.dptr(v),
                                                L.dptr(w), *geom._geo(), L.dptr(geom.vol_eff)))
    phi = DeviceArray((g.Nx, g.Ny, g.Nz))
    solve_source_buffer(phi, solver)
    bk = {"periodic": L.PERIODIC, "neumann": L.NOFLUX, "noflux": L.NOFLUX}
    L.check(L.init().aeo_cutcell_project_correct(g.handle, u.precision, L.dptr(u), L.dptr(v), L.dptr(w), L.dptr(phi),
                                                 *geom._geo(), *[bk[b] for b in bcs], 1.0))
    return phi
