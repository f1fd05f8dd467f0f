"""Synthetic, seeded inputs for the advection workloads.

All generation uses numpy's PCG64 generator with the seed recorded by the
caller.  Fields are float64 arrays of shape (nz, ny, nx) with periodic
lateral ghosts filled (qnbc_p4 images: x(1:3) = x(nx-4:nx-2),
x(nx-1:nx) = x(4:5), same in y).

* ``cone_2d``: the 2D_ADV/CONE style case (src/IO/initial.f90:582-679
  describes the reference's analytic setups): cosine bell of peak 1 plus
  non-negative uniform noise, constant wind.
* ``turbulent_wind``: divergence-free C-grid wind, the discrete curl of a
  seeded vector potential with a |k|^(-17/6) amplitude spectrum (shell
  kinetic-energy spectrum ~ k^(-5/3)), horizontal potential components
  vanishing at the floor and lid so w = 0 there, scaled to a target RMS.
* ``lognormal`` / ``theta_profile``: positive-definite scalar and potential
  temperature 300 K + 3 K/km + noise.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid


def fill_periodic(g: Grid, x: np.ndarray) -> np.ndarray:
    """Periodic lateral ghost images (qnbc_p4, bc.f90:409-436), in place."""
    nx, ny = g.nx, g.ny
    x[:, :, 0:3] = x[:, :, nx - 5:nx - 2]
    x[:, :, nx - 2:nx] = x[:, :, 3:5]
    if g.three_d:
        x[:, 0:3, :] = x[:, ny - 5:ny - 2, :]
        x[:, ny - 2:ny, :] = x[:, 3:5, :]
    return x


def embed(g: Grid, interior: np.ndarray) -> np.ndarray:
    """Full array from an interior (nz, MAXY or 1, MAXX) block, ghosts filled."""
    a = np.zeros(g.shape)
    a[g.interior()] = interior
    return fill_periodic(g, a)


def cone_2d(g: Grid, seed=1, radius_cells=15.0, noise=1.0e-3, u0=10.0, w0=2.0):
    """Cosine-bell cone (peak 1) plus U[0, noise) noise; u = u0 on every x
    face, w = w0 on every z face above the floor (w(1) = 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mx, nz = g.maxx, g.nz
    xc = (np.arange(mx) + 0.5)
    zc = (np.arange(nz) + 0.5)
    X, Z = np.meshgrid(xc, zc)
    r = np.sqrt((X - mx / 2) ** 2 + (Z - nz / 2) ** 2)
    q = np.where(r < radius_cells, 0.5 * (1.0 + np.cos(np.pi * r / radius_cells)), 0.0)
    q = q + rng.uniform(0.0, noise, size=q.shape)
    qf = embed(g, q[:, None, :])
    u = np.full(g.shape, u0)
    v = np.zeros(g.shape)
    w = np.full(g.shape, w0)
    w[0] = 0.0
    return qf, u, v, w


def _spectral_field(rng, shape, slope, dz_periodic=True):
    """Real random field on a periodic grid with amplitude |k|^slope."""
    kz = np.fft.fftfreq(shape[0])
    ky = np.fft.fftfreq(shape[1])
    kx = np.fft.rfftfreq(shape[2])
    K = np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    K[0, 0, 0] = np.inf
    amp = K ** slope
    ph = rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape)
    return np.fft.irfftn(amp * ph, s=shape, axes=(0, 1, 2))


def turbulent_wind(g: Grid, seed=2, rms=2.0):
    """Divergence-free C-grid velocity (u on west faces, v on south faces, w
    on bottom faces).  Ax lives at (i, j-1/2, k-1/2), Ay at (i-1/2, j, k-1/2),
    Az at (i-1/2, j-1/2, k); u = dAz/dy - dAy/dz, v = dAx/dz - dAz/dx,
    w = dAy/dx - dAx/dy, so the discrete divergence vanishes identically."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mx, my, nz = g.maxx, max(g.maxy, 1), g.nz
    nzp = 2 * (nz + 1)
    slope = -17.0 / 6.0
    env = np.sin(np.pi * np.arange(nz + 1) / nz)
    env[0] = env[nz] = 0.0  # floor and lid faces
    env = env[:, None, None]
    Ax = _spectral_field(rng, (nzp, my, mx), slope)[: nz + 1] * env     # levels k-1/2, k=1..nz+1
    Ay = _spectral_field(rng, (nzp, my, mx), slope)[: nz + 1] * env
    Az = _spectral_field(rng, (nzp, my, mx), slope)[:nz]                # levels k
    dzw = g.dzw[:, None, None]
    roll = lambda a, s, ax: np.roll(a, s, axis=ax)
    # u(i-1/2, j, k): Az(i-1/2, j+1/2) - Az(i-1/2, j-1/2)  -> Az[j+1]-Az[j]
    if g.three_d:
        u = (roll(Az, -1, 1) - Az) / g.dy - (Ay[1:] - Ay[:-1]) / dzw
        v = (Ax[1:] - Ax[:-1]) / dzw - (roll(Az, -1, 2) - Az) / g.dx
        w = (roll(Ay, -1, 2) - Ay) / g.dx - (roll(Ax, -1, 1) - Ax) / g.dy
    else:
        u = -(Ay[1:] - Ay[:-1]) / dzw
        v = np.zeros_like(u)
        w = (roll(Ay, -1, 2) - Ay) / g.dx
    w = w[:nz]  # bottom faces k-1/2 for k=1..nz; w[0] = 0 from the envelope
    w[0] = 0.0
    cur = np.sqrt((np.mean(u ** 2) + np.mean(v ** 2) + np.mean(w ** 2)) / (3 if g.three_d else 2))
    s = rms / cur
    return embed(g, u * s), embed(g, v * s), embed(g, w * s)


def lognormal(g: Grid, seed=3, mu=-9.0, sigma=1.0, zero_frac=0.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    my = max(g.maxy, 1)
    q = np.exp(rng.normal(mu, sigma, size=(g.nz, my, g.maxx)))
    if zero_frac > 0:
        q[rng.random(q.shape) < zero_frac] = 0.0
    return embed(g, q)


def theta_profile(g: Grid, seed=4, noise=0.1):
    rng = np.random.Generator(np.random.PCG64(seed))
    my = max(g.maxy, 1)
    th = 300.0 + 0.003 * g.zc()[:, None, None] + noise * rng.standard_normal((g.nz, my, g.maxx))
    return embed(g, th)


def cfl_dt(g: Grid, u, v, w, cfl=0.5):
    """Largest dt with sum of directional Courant numbers <= cfl."""
    s = np.abs(u) / g.dx + np.abs(w) / g.dzw[:, None, None]
    if g.three_d:
        s = s + np.abs(v) / g.dy
    return float(cfl / s.max())


def config3_fields(g: Grid, u, v, w, seed=5, nspec=15):
    """The 20 prognostic fields of the fused multi-field batch, in batch
    order: the momenta rho0*(u, v, w) (FP32), theta and qv (FP64), then
    nspec positive-definite species (FP32; log-normal with mu from -20 to
    -6, sigma 1.5, 30 % of cells exactly zero).  rho0 = 1.2 exp(-z/8000) at
    cell centres.  Returns (fields as float64 arrays holding the values the
    batch stores, precision codes 'f32'/'f64')."""
    den0 = 1.2 * np.exp(-g.zc() / 8000.0)[:, None, None]
    r32 = lambda a: a.astype(np.float32).astype(np.float64)
    out = [r32(r32(a) * r32(den0)) for a in (u, v, w)]
    prec = ["f32"] * 3
    out += [theta_profile(g, seed=seed), lognormal(g, seed=seed + 1, mu=-5.0, sigma=0.3)]
    prec += ["f64", "f64"]
    for s in range(nspec):
        q = lognormal(g, seed=seed + 5 + s, mu=-20 + 14 * s / max(1, nspec - 1), sigma=1.5, zero_frac=0.3)
        out.append(r32(q))
        prec.append("f32")
    return out, prec


def config4_draw(seed=4, ndom=48, weno="weno"):
    """Seeded draw of the uneven-subdomain batch: per subdomain MAXX, MAXY in
    [48, 192], nz in [48, 96], scheme `weno` ('weno' or 'weno5') or 'muscl'
    (even odds), 4-16 species, wind RMS in [1, 4] m/s, and per species the
    log-normal mu in [-20, -6] (sigma 1.5).  Returns a list of dicts; field
    seeds are 1000 + d (wind) and 2000 + 20 d + s (species)."""
    rng = np.random.default_rng(seed)
    doms = []
    for d in range(ndom):
        mx, my = (int(x) for x in rng.integers(48, 193, 2))
        nz = int(rng.integers(48, 97))
        scheme = weno if rng.random() < 0.5 else "muscl"
        ns = int(rng.integers(4, 17))
        rms = float(rng.uniform(1.0, 4.0))
        mus = [float(rng.uniform(-20, -6)) for _ in range(ns)]
        doms.append({"maxx": mx, "maxy": my, "nz": nz, "scheme": scheme, "nspec": ns, "rms": rms, "mu": mus,
                     "wind_seed": 1000 + d, "species_seeds": [2000 + 20 * d + s for s in range(ns)]})
    return doms
