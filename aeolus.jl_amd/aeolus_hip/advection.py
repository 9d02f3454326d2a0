"""Advection operators: the AdvectionSchemes operator surface.

`advection_scheme(name)` returns a callable `scheme(tend, u, v, w, phi)` that
ADDS -∂(uⱼϕ)/∂xⱼ to `tend` (src/AdvectionSchemes.jl:1704-1775), and
`momentum_advect_weno_z(tu, tv, tw, u, v, w)` ADDS -∂(uⱼuᵢ)/∂xⱼ
(src/AdvectionSchemes.jl:1195-1270). Arguments are DeviceArrays of the grid's
storage shape; halos must be filled.
"""
from . import _lib as L

SCHEMES = {
    "centered": (L.CENTERED, 1),
    "upwind": (L.UPWIND, 1),
    "weno5": (L.WENO5, 3),
    "weno_js5": (L.WENO_JS5, 3),
    "weno_z5": (L.WENO_Z5, 3),
    "weno_z7": (L.WENO_Z7, 4),
}


def scheme_id(name):
    try:
        return SCHEMES[name][0]
    except KeyError:
        raise ValueError(f"unknown advection scheme {name!r}; expected one of {', '.join(SCHEMES)}") from None


class AdvectionScheme:
    """Callable wrapper of one scheme (src/AdvectionSchemes.jl:1688-1700)."""

    def __init__(self, name, grid, p=2):
        if name not in SCHEMES:
            raise ValueError(f"unknown advection scheme {name!r}; expected one of {', '.join(SCHEMES)}")
        if p not in (1, 2):
            raise ValueError(f"WENO-Z τ-exponent p must be 1 or 2 (got {p})")
        self.name = name
        self.grid = grid
        self.halo = SCHEMES[name][1]
        self.p = p

    def __call__(self, tend, u, v, w, phi):
        L.check(L.init().aeo_advect_scalar(self.grid.handle, tend.precision, SCHEMES[self.name][0], self.p,
                                           L.dptr(tend), L.dptr(u), L.dptr(v), L.dptr(w), L.dptr(phi)))
        return tend

    def momentum(self, tu, tv, tw, u, v, w):
        L.check(L.init().aeo_advect_momentum(self.grid.handle, tu.precision, SCHEMES[self.name][0], self.p,
                                             L.dptr(tu), L.dptr(tv), L.dptr(tw), L.dptr(u), L.dptr(v), L.dptr(w)))
        return tu, tv, tw

    def __repr__(self):
        return f"AdvectionScheme(:{self.name}, halo ≥ {self.halo})"


def advection_scheme(name, grid, p=2):
    """advection_scheme(name; p) (src/AdvectionSchemes.jl:1746-1775)."""
    return AdvectionScheme(name, grid, p)


def required_halo(s):
    return s.halo


def advect_weno_z(grid, tend, u, v, w, phi, order=5, p=2):
    """advect_weno_z!(tend, u, v, w, ϕ; order, p) (src/AdvectionSchemes.jl:731-735)."""
    if order not in (5, 7):
        raise ValueError(f"WENO-Z order must be 5 or 7 (got {order})")
    return AdvectionScheme("weno_z5" if order == 5 else "weno_z7", grid, p)(tend, u, v, w, phi)


def momentum_advect_weno_z(grid, tu, tv, tw, u, v, w, order=5, p=2):
    """momentum_advect_weno_z!(tu, tv, tw, u, v, w; order, p) (src/AdvectionSchemes.jl:1195-1200)."""
    if order not in (5, 7):
        raise ValueError(f"WENO-Z order must be 5 or 7 (got {order})")
    return AdvectionScheme("weno_z5" if order == 5 else "weno_z7", grid, p).momentum(tu, tv, tw, u, v, w)
