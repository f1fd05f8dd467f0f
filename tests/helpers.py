"""Shared helpers of the test-suite (grids, seeded fields, norms)."""
import numpy as np

from fdd3260_amd import cases
from fdd3260_amd import grid as gm


def case3d(maxx=26, maxy=22, nz=18, seed=5, bcl=("per", "per", "per", "per", "nnf", "nnf"), zero_frac=0.0):
    g = gm.stretched(maxx, maxy, nz, bcl=bcl)
    u, v, w = cases.turbulent_wind(g, seed=seed, rms=2.0)
    q = cases.lognormal(g, seed=seed + 1, zero_frac=zero_frac)
    return g, u, v, w, q


def case2d(maxx=40, nz=20, seed=7, bcl=("per", "per", "per", "per", "nnf", "nnf")):
    g = gm.stretched(maxx, 0, nz, bcl=bcl)
    u, v, w = cases.turbulent_wind(g, seed=seed, rms=2.0)
    q = cases.lognormal(g, seed=seed + 1)
    return g, u, v, w, q


def relmax(got, ref):
    """Norm-wise relative error max|got-ref| / max|ref|."""
    den = max(float(np.abs(ref).max()), 1e-300)
    return float(np.abs(got - ref).max()) / den


def split_flips(got, ref, g, thr=1e-4):
    """Per-cell error relative to the field's interior RMS, split at thr:
    (number of cells above thr, RMS of the cells at or below thr).  The
    cells above are those where FP32 storage of a stage value flipped the
    ratio-20 smoothness switch (a flip moves one cell by 1e-3..1e-1 of the
    field RMS); the rest carry FP32 round-off."""
    I = g.interior()
    e = np.abs(got - ref)[I] / max(float(np.sqrt(np.mean(ref[I] ** 2))), 1e-300)
    out = e > thr
    return int(out.sum()), float(np.sqrt(np.mean(np.where(out, 0.0, e) ** 2)))
