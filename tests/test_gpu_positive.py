"""mimadv_advance with MIMADV_POSITIVE (the positivity-preserving stages;
no reference counterpart) against the oracle's orc_positive_stage, bit for
bit, and its invariants at the plain benchmark's size:

* SSP-RK3 and Heun (Shu-Osher form) steps on 3D and 2D grids, zero-laden
  log-normal data at Courant number 0.9, where the scaling is active:
  equal to the oracle's positivity steps;
* smooth positive data at a small time step: no cell needs scaling, and
  the steps equal the plain MIMADV_INT_SSPRK3 steps;
* 256x256x128 (the plain benchmark's grid and wind), log-normal scalar
  with 30 % zeros, 100 SSP-RK3 steps at Courant number 0.9: the plain path
  goes negative, the positivity path stays >= 0 and conserves the
  dzw-weighted mass to round-off;
* refusals: Euler, other schemes, advective form, an open floor.
"""
import numpy as np
import pytest
import torch

from fdd3260_amd import Advection, _lib, cases
from fdd3260_amd import grid as gm
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _case(three_d, seed, cfl=0.9, zero_frac=0.3, mu=-9.0, sigma=1.5):
    g = gm.stretched(34, 27 if three_d else 0, 19)
    u, v, w = cases.turbulent_wind(g, seed=seed, rms=3.0)
    if not three_d:
        v = np.zeros(g.shape)
    q = cases.lognormal(g, seed=seed + 1, mu=mu, sigma=sigma, zero_frac=zero_frac)
    dt = cases.cfl_dt(g, u, v, w, cfl)
    return g, u, v, w, q, dt


@pytest.mark.parametrize("integrator", ["ssprk3", "heun"])
@pytest.mark.parametrize("three_d", [True, False])
def test_positive_steps_match_oracle(cuda, integrator, three_d):
    print(f"case positive {integrator} 3d={three_d}", flush=True)
    g, u, v, w, q, dt = _case(three_d, seed=91)
    ora = Oracle(g)
    ref = q.copy()
    nsteps = 3
    stages = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0)) if integrator == "ssprk3" else ((0.0, 1.0), (0.5, 0.5))
    for _ in range(nsteps):
        q0, cur = ref.copy(), ref.copy()
        for c0, c1 in stages:
            cur = ora.positive_stage(dt, u, v, w, cur, q0, c0, c1)
        ref = cur
    code = _lib.INT_SSPRK3 if integrator == "ssprk3" else _lib.INT_HEUN
    with Advection(g) as a:
        U, V, W = (torch.from_numpy(x).to(cuda) for x in (u, v, w))
        t = torch.from_numpy(q).to(cuda)
        a.advance(code, nsteps, dt, U, V, W, [t], positive=True)
        a.stream_sync()
        got = t.cpu().numpy()
    assert np.array_equal(got, ref)
    assert np.all(got[g.interior()] >= 0.0)


def test_positive_equals_plain_where_nothing_is_scaled(cuda):
    print("case positive identity", flush=True)
    g, u, v, w, q, dt = _case(True, seed=93, cfl=0.05, zero_frac=0.0, mu=-3.0, sigma=0.1)
    with Advection(g) as a:
        U, V, W = (torch.from_numpy(x).to(cuda) for x in (u, v, w))
        p = torch.from_numpy(q).to(cuda)
        r = torch.from_numpy(q).to(cuda)
        a.advance(_lib.INT_SSPRK3, 4, dt, U, V, W, [p], positive=True)
        a.advance(_lib.INT_SSPRK3, 4, dt, U, V, W, [r])
        a.stream_sync()
        assert torch.equal(p, r)


def test_positive_at_benchmark_size(cuda):
    """100 SSP-RK3 steps at 256x256x128 and Courant number 0.9."""
    print("case positive 256x256x128, 100 steps", flush=True)
    g = gm.stretched(256, 256, 128)
    u, v, w = cases.turbulent_wind(g, seed=2, rms=2.0)
    q = cases.lognormal(g, seed=3, mu=-9.0, sigma=1.0, zero_frac=0.3)
    dt = cases.cfl_dt(g, u, v, w, 0.9)
    I = g.interior()
    dzw = torch.from_numpy(g.dzw).to(cuda)[:, None, None]
    with Advection(g) as a:
        U, V, W = (torch.from_numpy(x).to(cuda) for x in (u, v, w))
        p = torch.from_numpy(q).to(cuda)
        r = torch.from_numpy(q).to(cuda)
        m0 = float(torch.sum(p[I] * dzw))
        a.advance(_lib.INT_SSPRK3, 100, dt, U, V, W, [r])
        a.stream_sync()
        a.advance(_lib.INT_SSPRK3, 100, dt, U, V, W, [p], positive=True)
        a.stream_sync()
        ms = a.last_kernel_ms()
        pmin, rmin = float(p[I].min()), float(r[I].min())
        m1 = float(torch.sum(p[I] * dzw))
    print(f"  plain min {rmin:.3e}, positive min {pmin:.3e}, mass drift {abs(m1 - m0) / m0:.2e}, "
          f"{ms:.1f} ms for 100 steps", flush=True)
    assert rmin < 0.0
    assert pmin >= 0.0
    assert abs(m1 - m0) <= 1e-12 * m0


def test_positive_refusals(cuda):
    g, u, v, w, q, dt = _case(True, seed=95)
    with Advection(g) as a:
        U, V, W = (torch.from_numpy(x).to(cuda) for x in (u, v, w))
        t = torch.from_numpy(q).to(cuda)
        with pytest.raises(_lib.MimadvError, match="SSPRK3 or MIMADV_INT_HEUN"):
            a.advance(_lib.INT_EULER, 1, dt, U, V, W, [t], positive=True)
        with pytest.raises(_lib.MimadvError, match="flux-form"):
            a.advance(_lib.INT_SSPRK3, 1, dt, U, V, W, [t], adv=[True], positive=True)
        with pytest.raises(_lib.MimadvError, match="limited 'weno'"):
            a.advance(_lib.INT_SSPRK3, 1, dt, U, V, W, [t], limit=False, positive=True)
    with Advection(g, scal_adv="muscl") as a:
        U, V, W = (torch.from_numpy(x).to(cuda) for x in (u, v, w))
        t = torch.from_numpy(q).to(cuda)
        with pytest.raises(_lib.MimadvError, match="limited 'weno'"):
            a.advance(_lib.INT_SSPRK3, 1, dt, U, V, W, [t], positive=True)
    g2 = gm.stretched(34, 27, 19, bcl=("per", "per", "per", "per", "ope", "nnf"))
    with Advection(g2) as a:
        U, V, W = (torch.from_numpy(x).to(cuda) for x in (u, v, w))
        t = torch.from_numpy(q).to(cuda)
        with pytest.raises(_lib.MimadvError, match="'nnf' floor and lid"):
            a.advance(_lib.INT_SSPRK3, 1, dt, U, V, W, [t], positive=True)
