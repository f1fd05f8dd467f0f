"""Many independent periodic subdomains advanced by one launch per stage
(ctypes mirror of the mimadv_multi_* entry points of include/mimadv.h)."""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import MimadvError


class MultiDomain:
    """A batch of subdomains.  advs[d] is the Advection context of
    subdomain d (its grid and scheme), vel[d] = (u, v, w) its FP32 velocity
    and fields[d] its FP32 fields (contiguous GPU tensors of the grid's
    shape).  The tensors must stay alive as long as the batch."""

    def __init__(self, advs, vel, fields):
        self.lib = _lib.load()
        n = len(advs)
        if len(vel) != n or len(fields) != n:
            raise MimadvError("one velocity and one field list per subdomain")
        self.n = n
        self._keep = []
        ctxs = (C.c_void_p * n)(*[a._h.value for a in advs])
        us, vs, ws = ((C.c_void_p * n)() for _ in range(3))
        for d, (a, (u, v, w)) in enumerate(zip(advs, vel)):
            pu, pv, pw = a._vel32(u, v, w)
            us[d], vs[d], ws[d] = pu, pv, pw
        nf = (C.c_int * n)(*[len(f) for f in fields])
        qlists = []
        for d, fl in enumerate(fields):
            pr, pq = advs[d]._batch_fields(fl)
            if any(p != _lib.F32 for p in list(pr)[:len(fl)]):
                raise MimadvError("subdomain fields must be float32")
            qlists.append(pq)
        qq = (C.POINTER(C.c_void_p) * n)(*[C.cast(x, C.POINTER(C.c_void_p)) for x in qlists])
        self._keep += [ctxs, us, vs, ws, nf, qlists, qq, advs, vel, fields]
        h = C.c_void_p()
        rc = self.lib.mimadv_multi_create(n, ctxs, us, vs, ws, nf, qq, C.byref(h))
        self._h = h
        if rc != _lib.OK:
            msg = self.lib.mimadv_multi_last_error(h).decode() if h.value else "mimadv_multi_create failed"
            self.close()
            raise MimadvError(f"mimadv_multi_create: {msg} (status {rc})")

    def _check(self, rc, what):
        if rc != _lib.OK:
            raise MimadvError(f"{what}: {self.lib.mimadv_multi_last_error(self._h).decode()} (status {rc})")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.mimadv_multi_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tiles(self, tile_x=None, kchunk=None):
        tx = None if tile_x is None else (C.c_int * self.n)(*tile_x)
        kc = None if kchunk is None else (C.c_int * self.n)(*kchunk)
        self._check(self.lib.mimadv_multi_set_tiles(self._h, tx, kc), "mimadv_multi_set_tiles")

    def tiles(self):
        tx, ty, kc = ((C.c_int * self.n)() for _ in range(3))
        self._check(self.lib.mimadv_multi_get_tiles(self._h, tx, ty, kc), "mimadv_multi_get_tiles")
        return list(zip(tx, ty, kc))

    def autotune(self, reps=3):
        self._check(self.lib.mimadv_multi_autotune(self._h, int(reps)), "mimadv_multi_autotune")

    def advance(self, integrator, nsteps, dt):
        dts = (C.c_double * self.n)(*[float(x) for x in dt])
        self._check(self.lib.mimadv_multi_advance(self._h, int(integrator), int(nsteps), dts),
                    "mimadv_multi_advance")

    def last_ms(self):
        v = C.c_double()
        self._check(self.lib.mimadv_multi_last_ms(self._h, C.byref(v)), "mimadv_multi_last_ms")
        return v.value
