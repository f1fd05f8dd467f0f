"""Grid description of the MIMICA advection operator.

Mirrors the reference's compile-time layout and vertical metrics:
  * extents nx = MAXX+5, ny = MAXY+5 (1 in 2D), nz = MAXZ
    (INCLUDE/wgrid.csh:17-23), 3 west/south and 2 east/north ghosts;
  * dzw(k) = dz*fdz(k), dz1w(k) = 1/(dz*fdz(k)),
    dz1(k) = 1/(dz*0.5*(fdz(k)+fdz(k+1))), dz1(nz) = dz1(nz-1)
    (src/TOOLS/define_grid.f90:179-187);
  * fdz from a level file: fdz(1) = 2 zz(1)/dz,
    fdz(k) = 2 (zz(k)-zz(k-1))/dz - fdz(k-1) (define_grid.f90:167-170).
Arrays are held as C-ordered numpy arrays of shape (nz, ny, nx), whose
memory is the reference's Fortran (nx, ny, nz) column-major layout.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Grid:
    maxx: int
    maxy: int  # 0 for a 2D run
    nz: int
    dx: float
    dy: float
    dz: float
    fdz: np.ndarray
    bcl: tuple = ("per", "per", "per", "per", "nnf", "nnf")
    dzw: np.ndarray = field(init=False)
    dz1w: np.ndarray = field(init=False)
    dz1: np.ndarray = field(init=False)

    def __post_init__(self):
        f = np.ascontiguousarray(self.fdz, dtype=np.float64)
        if f.shape != (self.nz,):
            raise ValueError("fdz must have nz entries")
        self.fdz = f
        self.dzw = self.dz * f
        self.dz1w = 1.0 / (self.dz * f)
        dz1 = np.empty(self.nz)
        dz1[:-1] = 1.0 / ((self.dz * 0.5) * (f[:-1] + f[1:]))
        dz1[-1] = dz1[-2]
        self.dz1 = dz1

    @property
    def nx(self) -> int:
        return self.maxx + 5

    @property
    def ny(self) -> int:
        return self.maxy + 5 if self.maxy > 0 else 1

    @property
    def three_d(self) -> bool:
        return self.maxy > 0

    @property
    def shape(self) -> tuple:
        return (self.nz, self.ny, self.nx)

    def bc_codes(self):
        m = {"per": _lib.BC_PER, "nnf": _lib.BC_NNF, "ope": _lib.BC_OPE}
        return [m[b] for b in self.bcl]

    def interior(self):
        """Slices of the interior columns (all levels) in (k, j, i) order."""
        js = slice(3, self.ny - 2) if self.three_d else slice(0, 1)
        return (slice(None), js, slice(3, self.nx - 2))

    def zw(self) -> np.ndarray:
        """Height of the bottom face of every level."""
        return np.concatenate([[0.0], np.cumsum(self.dzw)[:-1]])

    def zc(self) -> np.ndarray:
        """Height of every level centre."""
        return self.zw() + 0.5 * self.dzw

    def to_c(self):
        """ctypes mimadv_grid; keeps the metric arrays alive on the object."""
        g = _lib.Grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.dx, g.dy, g.dz = self.dx, self.dy, self.dz
        dp = C.POINTER(C.c_double)
        g.fdz0 = self.fdz.ctypes.data_as(dp)
        g.dzw = self.dzw.ctypes.data_as(dp)
        g.dz1w = self.dz1w.ctypes.data_as(dp)
        g.dz1 = self.dz1.ctypes.data_as(dp)
        for n, b in enumerate(self.bc_codes()):
            g.bcl[n] = b
        return g


def fdz_from_levels(zz: np.ndarray, dz: float) -> np.ndarray:
    """fdz from level-centre heights as read from a grid file
    (define_grid.f90:167-170)."""
    zz = np.asarray(zz, dtype=np.float64)
    f = np.empty_like(zz)
    f[0] = 2.0 * zz[0] / dz
    for k in range(1, zz.size):
        f[k] = 2.0 * (zz[k] - zz[k - 1]) / dz - f[k - 1]
    return f


def uniform(maxx, maxy, nz, dx=50.0, dy=50.0, dz=50.0, bcl=("per", "per", "per", "per", "nnf", "nnf")) -> Grid:
    return Grid(maxx, maxy, nz, dx, dy, dz, np.ones(nz), tuple(bcl))


def stretched(maxx, maxy, nz, dx=50.0, dy=50.0, dz_bottom=20.0, dz_top=80.0,
              bcl=("per", "per", "per", "per", "nnf", "nnf")) -> Grid:
    """Cell thickness growing linearly from dz_bottom to dz_top; the
    reference dz is dz_bottom and fdz(k) = thickness(k)/dz."""
    t = dz_bottom + (dz_top - dz_bottom) * np.arange(nz) / max(nz - 1, 1)
    return Grid(maxx, maxy, nz, dx, dy, dz_bottom, t / dz_bottom, tuple(bcl))
