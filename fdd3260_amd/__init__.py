"""MI355X-native advection operator of the MIMICA LES model.

libmimadv.so (HIP, gfx950) implements the reference's advection operator
(src/TOOLS/advection.f90, advq_{weno,ppm,muscl}.f90, src/INTEG/windadv.f90)
behind a C ABI (include/mimadv.h).  This package holds the host-side mirror
of the reference interface (``Advection``), the grid description and the
synthetic-case builders.
"""
from ._lib import MimadvError, load  # noqa: F401
from .advection import Advection  # noqa: F401
from .grid import Grid, stretched, uniform  # noqa: F401
from .multidomain import MultiDomain  # noqa: F401

__all__ = ["Advection", "Grid", "MimadvError", "MultiDomain", "load", "stretched", "uniform"]
