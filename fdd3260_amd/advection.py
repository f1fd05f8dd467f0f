"""Host-side mirror of MIMICA's advection interface over libmimadv.

The reference exposes (src/TOOLS/advection.f90:79)
    advect, advection_x, advection_y, advection_z, advection_ac, ...
and (src/INTEG/windadv.f90:41) qnadv_uvw.  ``Advection`` keeps those names
and argument meanings for arrays held in Python:

  * torch tensors on the GPU are passed as device pointers;
  * numpy arrays (host memory) are staged by the library
    (MIMADV_HOST_PTRS), like the Fortran shim does for a CPU caller.

Arrays are float64 of shape (nz, ny, nx), i.e. the reference's Fortran
(nx, ny, nz) layout.  Errors raise ``MimadvError`` carrying the library's
message (the reference has no error returns; it stops later in checkall).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MimadvError
from .grid import Grid

SCHEMES = {"weno": _lib.WENO, "ppm": _lib.PPM, "muscl": _lib.MUSCL, "weno5": _lib.WENO5, "quick": _lib.QUICK,
           "lw": _lib.LW}
LIMITERS = {"minmod": _lib.LIM_MINMOD, "MC": _lib.LIM_MC, "VL": _lib.LIM_VL, "Fromm": _lib.LIM_FROMM}


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


class Advection:
    """One library context: grid + options (scal_adv, muscllimiter,
    ppmlimiter, lim_tol, mom_ord, lavisc) + a HIP stream."""

    def __init__(self, grid: Grid, scal_adv="weno", muscllimiter="minmod", ppmlimiter="PPM",
                 lim_tol=1.0e-9, mom_ord=4, lavisc=True, weno_bottom=_lib.WENO_BOTTOM_X0_IS_X1):
        self.lib = _lib.load()
        self.grid = grid
        self._cgrid = grid.to_c()
        self.opts = _lib.Opts()
        self._set_opts(scal_adv, muscllimiter, ppmlimiter, lim_tol, mom_ord, lavisc, weno_bottom)
        h = C.c_void_p()
        rc = self.lib.mimadv_create(C.byref(self._cgrid), C.byref(self.opts), C.byref(h))
        self._h = h
        if rc != _lib.OK:
            msg = self.lib.mimadv_last_error(h).decode() if h.value else "mimadv_create failed"
            if h.value:
                self.lib.mimadv_destroy(h)
            self._h = C.c_void_p()
            raise MimadvError(f"mimadv_create: {msg} (status {rc})")

    # -- options -----------------------------------------------------------
    def _set_opts(self, scal_adv, muscllimiter, ppmlimiter, lim_tol, mom_ord, lavisc, weno_bottom):
        if scal_adv not in SCHEMES:
            raise MimadvError(f"unknown scal_adv {scal_adv!r}")
        if muscllimiter not in LIMITERS:
            raise MimadvError(f"unknown muscllimiter {muscllimiter!r}")
        o = self.opts
        o.scheme = SCHEMES[scal_adv]
        o.muscl_limiter = LIMITERS[muscllimiter]
        o.ppm_limiter = 1 if ppmlimiter == "PPM" else 0
        o.lim_tol = float(lim_tol)
        o.weno_bottom = int(weno_bottom)
        o.mom_ord = int(mom_ord)
        o.lavisc = 1 if lavisc else 0
        self.scal_adv = scal_adv

    def set_options(self, scal_adv=None, muscllimiter=None, ppmlimiter=None, lim_tol=None, mom_ord=None,
                    lavisc=None, weno_bottom=None):
        inv_s = {v: k for k, v in SCHEMES.items()}
        inv_l = {v: k for k, v in LIMITERS.items()}
        o = self.opts
        self._set_opts(
            scal_adv if scal_adv is not None else inv_s[o.scheme],
            muscllimiter if muscllimiter is not None else inv_l[o.muscl_limiter],
            ppmlimiter if ppmlimiter is not None else ("PPM" if o.ppm_limiter else "none"),
            lim_tol if lim_tol is not None else o.lim_tol,
            mom_ord if mom_ord is not None else o.mom_ord,
            lavisc if lavisc is not None else bool(o.lavisc),
            weno_bottom if weno_bottom is not None else o.weno_bottom,
        )
        self._check(self.lib.mimadv_set_opts(self._h, C.byref(self.opts)), "mimadv_set_opts")

    # -- plumbing ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.mimadv_destroy(self._h)
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

    def _check(self, rc, what):
        if rc != _lib.OK:
            raise MimadvError(f"{what}: {self.lib.mimadv_last_error(self._h).decode()} (status {rc})")

    def _ptrs(self, arrays, writable=()):
        """Pointers and the HOST_PTRS flag for a homogeneous list of arrays."""
        host = None
        ptrs = []
        shape = self.grid.shape
        for a in arrays:
            if a is None:
                ptrs.append(None)
                continue
            if _is_torch(a):
                import torch

                if a.dtype != torch.float64 or not a.is_contiguous() or tuple(a.shape) != shape:
                    raise MimadvError(f"expected a contiguous float64 tensor of shape {shape}")
                if not a.is_cuda:
                    raise MimadvError("torch tensors must live on the GPU (use numpy for host arrays)")
                is_host = False
                ptrs.append(a.data_ptr())
            else:
                if not isinstance(a, np.ndarray) or a.dtype != np.float64 or a.shape != shape \
                        or not a.flags.c_contiguous:
                    raise MimadvError(f"expected a C-contiguous float64 array of shape {shape}")
                is_host = True
                ptrs.append(a.ctypes.data)
            if host is None:
                host = is_host
            elif host != is_host:
                raise MimadvError("cannot mix host and device arrays in one call")
        for a in writable:
            if isinstance(a, np.ndarray) and not a.flags.writeable:
                raise MimadvError("output array is read-only")
        return ptrs, (_lib.HOST_PTRS if host else 0)

    def stream_sync(self):
        self._check(self.lib.mimadv_synchronize(self._h), "mimadv_synchronize")

    def set_stream(self, stream_ptr):
        """Run on a caller stream (a hipStream_t handle; 0 is the null stream, which is
        torch's default stream); _lib.OWN_STREAM returns to the context's own stream."""
        self._check(self.lib.mimadv_set_stream(self._h, stream_ptr), "mimadv_set_stream")

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        self._check(self.lib.mimadv_last_kernel_ms(self._h, C.byref(v)), "mimadv_last_kernel_ms")
        return v.value

    # -- scalar advection (advection.f90) -----------------------------------
    def scalar_tendency(self, dt0, u, v, w, q_list, xs_list, adv=None, dirs=_lib.DIR_ALL, overwrite=False,
                        limit=True, sorder=2):
        """xs_f = ((xs_f - X) - Y) - Z for every field (advection_x/y/z); limit and sorder select the line
        kernel as the reference does (unlimited: QUICK or Lax-Wendroff; sorder 1: Lax-Wendroff)."""
        nf = len(q_list)
        if len(xs_list) != nf:
            raise MimadvError("q and xs lists differ in length")
        ptrs, hflag = self._ptrs([u, v, w] + list(q_list) + list(xs_list), writable=xs_list)
        pu, pv, pw = ptrs[:3]
        pq = _lib.ptr_array(ptrs[3:3 + nf])
        px = _lib.ptr_array(ptrs[3 + nf:])
        ad = None
        if adv is not None:
            ad = (C.c_int * nf)(*[1 if a else 0 for a in adv])
        flags = (int(dirs) | hflag | (_lib.OVERWRITE if overwrite else 0) | (0 if limit else _lib.UNLIMITED)
                 | (_lib.SORDER1 if sorder == 1 else 0))
        rc = self.lib.mimadv_scalar_tendency(self._h, float(dt0), pu, pv, pw, nf, pq, px, ad, flags)
        self._check(rc, "mimadv_scalar_tendency")

    def advection_x(self, dtl, xtmp, xs, u, adv=False):
        self.scalar_tendency(dtl, u, None, None, [xtmp], [xs], [adv], dirs=_lib.DIR_X)

    def advection_y(self, dtl, xtmp, xs, v, adv=False):
        self.scalar_tendency(dtl, None, v, None, [xtmp], [xs], [adv], dirs=_lib.DIR_Y)

    def advection_z(self, dtl, xtmp, xs, w, adv=False):
        self.scalar_tendency(dtl, None, None, w, [xtmp], [xs], [adv], dirs=_lib.DIR_Z)

    def advection_ac(self, dt0, u, v, w, xx, xs, adv=False):
        """Single tracer: xs += xsx + xsy + xsz (advection.f90:801-869)."""
        self.scalar_tendency(dt0, u, v, w, [xx], [xs], [adv])

    # -- momentum (windadv.f90) --------------------------------------------
    def qnadv_uvw(self, ntime, mom, wind_adv, out):
        """out (ua, va, wa) = +div of the momenta mom advected by wind_adv."""
        mu, mv, mw = mom
        ut, vt, wt = wind_adv
        ua, va, wa = out
        ptrs, hflag = self._ptrs([mu, mv, mw, ut, vt, wt, ua, va, wa], writable=out)
        rc = self.lib.mimadv_momentum_tendency(self._h, int(ntime), *ptrs, hflag)
        self._check(rc, "mimadv_momentum_tendency")

    # -- boundaries, stage combine, budgets, stepping -----------------------
    def fill_halo(self, fields):
        ptrs, hflag = self._ptrs(list(fields), writable=fields)
        rc = self.lib.mimadv_fill_halo(self._h, len(fields), _lib.ptr_array(ptrs), hflag)
        self._check(rc, "mimadv_fill_halo")

    def fill_winds(self, u, v, w):
        """lbcu_p4: ghost fill of the winds (in place)."""
        ptrs, hflag = self._ptrs([u, v, w], writable=[u, v, w])
        self._check(self.lib.mimadv_fill_winds(self._h, *ptrs, hflag), "mimadv_fill_winds")

    def integ_mom(self, sorder, im, dt, z0, zmax, zdamp, sponge, old, a, tend, x):
        """integ_mom on (u, v, w) triples; tend and x updated in place."""
        ptrs, hflag = self._ptrs(list(old) + list(a) + list(tend) + list(x), writable=list(tend) + list(x))
        z = np.ascontiguousarray(z0, dtype=np.float64)
        zp = z.ctypes.data_as(C.POINTER(C.c_double))
        arr = [_lib.ptr_array(ptrs[3 * n:3 * n + 3]) for n in range(4)]
        rc = self.lib.mimadv_integ_mom(self._h, int(sorder), int(im), float(dt), zp, float(zmax), float(zdamp),
                                       int(sponge), *arr, hflag)
        self._check(rc, "mimadv_integ_mom")

    def integ_scal(self, sorder, im, dt, xold, a, tend, x):
        ptrs, hflag = self._ptrs([xold, a, tend, x], writable=[tend, x])
        arr = [_lib.ptr_array([p]) for p in ptrs]
        rc = self.lib.mimadv_integ_scal(self._h, int(sorder), int(im), float(dt), 1, *arr, hflag)
        self._check(rc, "mimadv_integ_scal")

    def totalav(self, x):
        (p,), hflag = self._ptrs([x])
        mean, wsum = C.c_double(), C.c_double()
        rc = self.lib.mimadv_totalav(self._h, p, C.byref(mean), C.byref(wsum), hflag)
        self._check(rc, "mimadv_totalav")
        return mean.value, wsum.value

    def advance(self, integrator, nsteps, dt, u, v, w, fields, adv=None, limit=True, positive=False):
        ptrs, hflag = self._ptrs([u, v, w] + list(fields), writable=fields)
        if hflag:
            raise MimadvError("advance works on device tensors")
        nf = len(fields)
        ad = None
        if adv is not None:
            ad = (C.c_int * nf)(*[1 if a else 0 for a in adv])
        rc = self.lib.mimadv_advance(self._h, int(integrator), int(nsteps), float(dt), ptrs[0], ptrs[1], ptrs[2],
                                     nf, _lib.ptr_array(ptrs[3:]), ad,
                                     (0 if limit else _lib.UNLIMITED) | (_lib.POSITIVE if positive else 0))
        self._check(rc, "mimadv_advance")

    # -- batched mixed-precision path ---------------------------------------
    def set_tuning(self, group32=0, kchunk=0, blocks_per_cu=0, tile_y=0, tile_x=0):
        t = _lib.Tuning(int(group32), int(kchunk), int(blocks_per_cu), int(tile_y), int(tile_x))
        self._check(self.lib.mimadv_set_tuning(self._h, C.byref(t)), "mimadv_set_tuning")

    def _batch_fields(self, fields):
        import torch

        shape = self.grid.shape
        prec, ptrs = [], []
        for f in fields:
            if not _is_torch(f) or not f.is_cuda or not f.is_contiguous() or tuple(f.shape) != shape:
                raise MimadvError(f"batched fields must be contiguous GPU tensors of shape {shape}")
            if f.dtype == torch.float32:
                prec.append(_lib.F32)
            elif f.dtype == torch.float64:
                prec.append(_lib.F64)
            else:
                raise MimadvError("batched fields must be float32 or float64")
            ptrs.append(f.data_ptr())
        return (C.c_int * max(1, len(prec)))(*prec), _lib.ptr_array(ptrs)

    def _vel32(self, u, v, w):
        import torch

        shape = self.grid.shape
        out = []
        for a in (u, v, w):
            if a is None:
                out.append(None)
                continue
            if not _is_torch(a) or a.dtype != torch.float32 or not a.is_cuda or not a.is_contiguous() \
                    or tuple(a.shape) != shape:
                raise MimadvError(f"batched velocities must be contiguous float32 GPU tensors of shape {shape}")
            out.append(a.data_ptr())
        return out

    def batch_tendency(self, u, v, w, fields, outs, dirs=_lib.DIR_ALL, momentum=False, flags=0):
        """momentum: fields[0:3] are the FP32 momenta rho0*(u, v, w), advanced with the momentum operator.
        flags: further MIMADV_* call flags, passed through."""
        pr, pq = self._batch_fields(fields)
        pr2, px = self._batch_fields(outs)
        if list(pr) != list(pr2):
            raise MimadvError("tendencies must have the precision of their fields")
        pu, pv, pw = self._vel32(u, v, w)
        flags = int(dirs) | (_lib.MOMENTUM if momentum else 0) | int(flags)
        rc = self.lib.mimadv_batch_tendency(self._h, pu, pv, pw, len(fields), pr, pq, px, flags)
        self._check(rc, "mimadv_batch_tendency")

    def batch_advance(self, integrator, nsteps, dt, u, v, w, fields, graph=False, momentum=False, flags=0,
                      profile=False):
        """profile: time every launch with events (eager launches); read with batch_kernel_ms()."""
        pr, pq = self._batch_fields(fields)
        pu, pv, pw = self._vel32(u, v, w)
        flags = ((_lib.GRAPH if graph else 0) | (_lib.MOMENTUM if momentum else 0) | (_lib.PROFILE if profile else 0)
                 | int(flags))
        rc = self.lib.mimadv_batch_advance(self._h, int(integrator), int(nsteps), float(dt), pu, pv, pw,
                                           len(fields), pr, pq, flags)
        self._check(rc, "mimadv_batch_advance")

    def batch_budget_device(self, fields, out):
        """Budgets of the fields into the float64 GPU tensor out[len(fields)]
        with no device allocation (mimadv_batch_budget_device)."""
        import torch

        pr, pq = self._batch_fields(fields)
        if not _is_torch(out) or out.dtype != torch.float64 or not out.is_cuda or out.numel() < len(fields):
            raise MimadvError("out must be a float64 GPU tensor with one entry per field")
        rc = self.lib.mimadv_batch_budget_device(self._h, len(fields), pr, pq, out.data_ptr())
        self._check(rc, "mimadv_batch_budget_device")

    def batch_kernel_ms(self):
        """Per-kind kernel time of the last profiled batched call: {kind: (ms, launches)}
        for kinds 'momentum', 'f64', 'f32'."""
        ms = (C.c_double * 3)()
        cnt = (C.c_int * 3)()
        self._check(self.lib.mimadv_batch_kernel_ms(self._h, ms, cnt), "mimadv_batch_kernel_ms")
        return {k: (ms[n], cnt[n]) for n, k in enumerate(("momentum", "f64", "f32"))}

    def batch_budget(self, fields):
        pr, pq = self._batch_fields(fields)
        out = (C.c_double * max(1, len(fields)))()
        self._check(self.lib.mimadv_batch_budget(self._h, len(fields), pr, pq, out), "mimadv_batch_budget")
        return [out[n] for n in range(len(fields))]

    # -- adjoint of the plain FP64 'weno' path --------------------------------
    def adjoint_tendency(self, u, v, w, q, lam, qbar, ubar, vbar, wbar):
        """qbar = (dL/dq)^T lam (interior cells); ubar, vbar, wbar += (dL/d(u,v,w))^T lam."""
        ptrs, hflag = self._ptrs([u, v, w, q, lam, qbar, ubar, vbar, wbar], writable=(qbar, ubar, vbar, wbar))
        if hflag:
            raise MimadvError("the adjoint works on device tensors")
        rc = self.lib.mimadv_adjoint_tendency(self._h, *ptrs)
        self._check(rc, "mimadv_adjoint_tendency")

    def adjoint_advance(self, integrator, nsteps, dt, u, v, w, q0, qbar, ubar, vbar, wbar, checkpoint=0):
        """Adjoint of advance(integrator, nsteps, dt, u, v, w, [q]) from q0: qbar holds dJ/dq(T) on entry and
        dJ/dq0 on return; ubar, vbar, wbar are accumulated."""
        ptrs, hflag = self._ptrs([u, v, w, q0, qbar, ubar, vbar, wbar], writable=(qbar, ubar, vbar, wbar))
        if hflag:
            raise MimadvError("the adjoint works on device tensors")
        rc = self.lib.mimadv_adjoint_advance(self._h, int(integrator), int(nsteps), float(dt), ptrs[0], ptrs[1],
                                             ptrs[2], ptrs[3], ptrs[4], ptrs[5], ptrs[6], ptrs[7], int(checkpoint))
        self._check(rc, "mimadv_adjoint_advance")

    def advance_record(self, integrator, nsteps, dt, u, v, w, q, checkpoint, ck):
        """advance() of one flux-form field that copies the state at steps 0, K, 2K, ... into ck."""
        ptrs, hflag = self._ptrs([u, v, w, q] + list(ck), writable=[q] + list(ck))
        if hflag:
            raise MimadvError("advance_record works on device tensors")
        rc = self.lib.mimadv_advance_record(self._h, int(integrator), int(nsteps), float(dt), ptrs[0], ptrs[1],
                                            ptrs[2], ptrs[3], int(checkpoint), _lib.ptr_array(ptrs[4:]))
        self._check(rc, "mimadv_advance_record")

    def adjoint_advance_recorded(self, integrator, nsteps, dt, u, v, w, checkpoint, ck, qbar, ubar, vbar, wbar):
        """Adjoint from checkpoints recorded by advance_record (no forward sweep of its own)."""
        ptrs, hflag = self._ptrs([u, v, w, qbar, ubar, vbar, wbar] + list(ck), writable=(qbar, ubar, vbar, wbar))
        if hflag:
            raise MimadvError("the adjoint works on device tensors")
        rc = self.lib.mimadv_adjoint_advance_recorded(self._h, int(integrator), int(nsteps), float(dt), ptrs[0],
                                                      ptrs[1], ptrs[2], int(checkpoint), _lib.ptr_array(ptrs[7:]),
                                                      ptrs[3], ptrs[4], ptrs[5], ptrs[6])
        self._check(rc, "mimadv_adjoint_advance_recorded")

    # -- density weighting (funcpack.f90, anelastic) ---------------------------
    def set_reference_density(self, den0, avden0=None):
        """den0(k) and (optionally) avden0(k); avden0 defaults to the fdz0-weighted face average."""
        d = np.ascontiguousarray(den0, dtype=np.float64)
        a = None if avden0 is None else np.ascontiguousarray(avden0, dtype=np.float64)
        if d.shape != (self.grid.nz,) or (a is not None and a.shape != d.shape):
            raise MimadvError("density profiles need nz values")
        dp = C.POINTER(C.c_double)
        rc = self.lib.mimadv_set_reference_density(self._h, d.ctypes.data_as(dp),
                                                   None if a is None else a.ctypes.data_as(dp))
        self._check(rc, "mimadv_set_reference_density")

    def get_momentum(self, u, v, w, inverse=False):
        """In place: u, v *= den0(k), w *= avden0(k) (inverse: get_velocities)."""
        ptrs, hflag = self._ptrs([u, v, w], writable=(u, v, w))
        rc = self.lib.mimadv_get_momentum(self._h, ptrs[0], ptrs[1], ptrs[2], 1 if inverse else 0, hflag)
        self._check(rc, "mimadv_get_momentum")

    def get_advective(self, u0shift, v0shift, u, v, w, ua, va, wa):
        ptrs, hflag = self._ptrs([u, v, w, ua, va, wa], writable=(ua, va, wa))
        rc = self.lib.mimadv_get_advective(self._h, float(u0shift), float(v0shift), *ptrs, hflag)
        self._check(rc, "mimadv_get_advective")

    def advection_1d(self, w0, q_list, xs_list, dens=None, den0=False):
        """advection_ac_1d: xs_f -= dref*w0*grad_1d(q_f/dref); dref = den0 profile (den0=True), the 3D
        array dens, or 1.  w0 is a host profile of nz values."""
        nf = len(q_list)
        if len(xs_list) != nf:
            raise MimadvError("q and xs lists differ in length")
        w = np.ascontiguousarray(w0, dtype=np.float64)
        if w.shape != (self.grid.nz,):
            raise MimadvError(f"w0 must hold nz = {self.grid.nz} values")
        arrs = list(q_list) + list(xs_list) + ([dens] if dens is not None else [])
        ptrs, hflag = self._ptrs(arrs, writable=xs_list)
        pd = ptrs[2 * nf] if dens is not None else None
        flags = hflag | (_lib.DREF_DEN0 if den0 else 0)
        rc = self.lib.mimadv_advection_1d(self._h, w.ctypes.data_as(C.POINTER(C.c_double)), pd, nf,
                                          _lib.ptr_array(ptrs[:nf]), _lib.ptr_array(ptrs[nf:2 * nf]), flags)
        self._check(rc, "mimadv_advection_1d")

    def get_conserved(self, dens, fields, inverse=False):
        """In place: q *= dens (inverse: q /= dens, get_scalars) for every field."""
        ptrs, hflag = self._ptrs([dens] + list(fields), writable=fields)
        rc = self.lib.mimadv_get_conserved(self._h, ptrs[0], len(fields), _lib.ptr_array(ptrs[1:]),
                                           1 if inverse else 0, hflag)
        self._check(rc, "mimadv_get_conserved")
