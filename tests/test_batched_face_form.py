"""The rearranged WENO face of the batched kernels (fused_faces.hpp,
weno_face) against the reference face (advq_weno.f90:92-110, weights
:329-350, limiter_weno :462-491), both restated here in numpy:

* in extended precision (64-bit mantissa) the two forms agree on random
  stencils with exact zeros: the rearrangement is an identity;
* on a periodic line of log-normal data with 30 % exact zeros, one SSP-RK3
  step with the rearranged face in FP32 (FP32 stage storage, FP64 flux
  divergence, as the batched kernels do) stays at FP32 round-off of the
  extended-precision reference, while the direct form b + phi (F - b) in
  FP32 is off by orders of magnitude more: the limiter pulls face values
  to -eps b / |minp - b + eps| (eps = 1e-15), and the next stage's phi
  depends on those 1e-15-sized cell values.

No GPU: this pins the algebra the kernel uses; the kernel itself is held to
the oracle by the GPU tests.
"""
import numpy as np
import pytest

LD = np.longdouble


def _ref_face(a, b, c, d, uu, t):
    """Reference face flux in dtype t."""
    a, b, c, d, uu = (x.astype(t) for x in (a, b, c, d, uu))
    g1, g2, eps, thr = t(1) / t(3), t(2) / t(3), t(1e-15), t(20)
    eta1, eta2, eta3 = (b - a) ** 2, (c - b) ** 2, (d - c) ** 2
    s1, s2, s3 = (eps + eta1) ** 2, (eps + eta2) ** 2, (eps + eta3) ** 2
    w1, w2, w3, w4 = g1 / s1, g2 / s2, g2 / s2, g1 / s3
    swl = np.maximum(eta1, eta2) / (np.minimum(eta1, eta2) + eps) > thr
    swr = np.maximum(eta2, eta3) / (np.minimum(eta2, eta3) + eps) > thr
    ww0 = np.where(swl, w1 / (w1 + w2), g1)
    ww1 = np.where(swl, w2 / (w1 + w2), g2)
    ww2 = np.where(swr, w3 / (w3 + w4), g2)
    ww3 = np.where(swr, w4 / (w3 + w4), g1)
    f1 = -(t(0.5) * a) + t(1.5) * b
    f2 = t(0.5) * b + t(0.5) * c
    f4 = t(1.5) * c - t(0.5) * d
    minp = np.minimum(np.minimum(f1, f2), f4)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.fmin(np.abs((t(0) - b) / (minp - b + eps)), t(1))
    f1, f2, f4 = (phi * (f - b) + b for f in (f1, f2, f4))
    fl = ww0 * f1 + ww1 * f2
    fr = ww2 * f2 + ww3 * f4
    return t(0.5) * fl * (uu + np.abs(uu)) + t(0.5) * fr * (uu - np.abs(uu))


def _weights(dm, dc, dp, uu, t):
    eps, thr = t(1e-15), t(20)
    pos = uu > 0
    dout = np.where(pos, dm, dp)
    ei, eo = dc * dc + eps, dout * dout + eps
    sw = np.maximum(eo, ei) > thr * np.minimum(eo, ei) + eps
    so, si = eo * eo, ei * ei
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        r = t(1) / (si + t(2) * so)
        wo = np.where(sw, si * r, t(1) / t(3))
        wi = np.where(sw, t(2) * so * r, t(2) / t(3))
    return pos, wo, wi


def _batched_face(a, b, c, d, uu, t, direct=False):
    """The batched kernels' face in dtype t; direct=True evaluates the
    limited value as b + phi (F - b) instead."""
    a, b, c, d, uu = (x.astype(t) for x in (a, b, c, d, uu))
    eps = t(1e-15)
    dm, dc, dp = b - a, c - b, d - c
    ffim1, ffip1, ffip2 = b + t(0.5) * dm, b + t(0.5) * dc, c - t(0.5) * dp
    minp = np.minimum(np.minimum(ffim1, ffip1), ffip2)
    pos, wo, wi = _weights(dm, dc, dp, uu, t)
    fo = np.where(pos, ffim1, ffip2)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        if direct:
            phi = np.fmin(np.abs(b * (t(1) / ((minp - b) + eps))), t(1))
            return uu * (b + phi * (t(0.5) * dc + wo * (fo - ffip1)))
        e = wi * (ffip1 - minp) + wo * (fo - minp)
        m = minp + eps
        tt = b * (eps - e) * (t(1) / (m - b))
        F = np.where((m < 0) != (b < 0), tt, np.where(np.abs(m) > t(2) * np.abs(b), t(2) * b - tt, minp + e))
    return uu * F


def _stencils(n, mu, seed, signed=False):
    rng = np.random.default_rng(seed)
    v = np.exp(rng.normal(mu, 1.5, (4, n)))
    if signed:
        # stage values of the reference can undershoot: mixed signs reach
        # the branch |minp + eps| > 2|b| of either sign
        v *= np.where(rng.random(v.shape) < 0.3, -1.0, 1.0)
    v[rng.random(v.shape) < 0.3] = 0.0
    v = v.astype(np.float32).astype(np.float64)
    uu = rng.normal(0.0, 1.0, n).astype(np.float32).astype(np.float64)
    return v, uu


@pytest.mark.parametrize("signed", [False, True])
@pytest.mark.parametrize("mu", [-20.0, -16.0, -12.0, -6.0])
def test_rearranged_face_is_the_reference_face(mu, signed):
    v, uu = _stencils(50000, mu, seed=int(-mu), signed=signed)
    ref = _ref_face(*v, uu, LD)
    got = _batched_face(*v, uu, LD)
    scale = np.abs(uu.astype(LD)) * (np.abs(v[1].astype(LD)) + np.abs(v[2].astype(LD))) + LD(1e-300)
    assert float(np.max(np.abs(got - ref) / scale)) <= 1e-15


def _ssprk3(q, u, dt, face, store):
    """One SSP-RK3 step on a periodic line; face i+1/2 uses u[i]; the flux
    divergence and the stage combine in FP64, then store() rounds."""
    q0, cur = q.copy(), q

    def tend(x):
        F = face(np.roll(x, 1), x, np.roll(x, -1), np.roll(x, -2), u).astype(np.float64)
        return -(F - np.roll(F, 1))

    for c0, c1 in ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0)):
        cur = store(c0 * q0 + c1 * (cur + dt * tend(cur)))
    return cur


def _ssprk3_ld(q, u, dt):
    q0, cur = q.astype(LD), q.astype(LD)
    ul, dtl = u.astype(LD), LD(dt)
    for c0, c1 in ((LD(0), LD(1)), (LD(3) / 4, LD(1) / 4), (LD(1) / 3, LD(2) / 3)):
        F = _ref_face(np.roll(cur, 1), cur, np.roll(cur, -1), np.roll(cur, -2), ul, LD)
        cur = c0 * q0 + c1 * (cur + dtl * (-(F - np.roll(F, 1))))
    return cur.astype(np.float64)


@pytest.mark.parametrize("mu", [-16.0, -12.0, -6.0])
def test_fp32_step_next_to_zeros_stays_at_round_off(mu):
    n = 40000
    rng = np.random.default_rng(int(-mu) + 100)
    q = np.exp(rng.normal(mu, 1.5, n))
    q[rng.random(n) < 0.3] = 0.0
    q = q.astype(np.float32).astype(np.float64)
    u = rng.normal(0.0, 1.0, n).astype(np.float32).astype(np.float64)
    dt = 0.5 / float(np.abs(u).max())
    truth = _ssprk3_ld(q, u, dt)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)

    def rel(x):
        return float(np.sqrt(np.mean((x - truth) ** 2)) / np.sqrt(np.mean(truth ** 2)))

    good = rel(_ssprk3(q, u, dt, lambda *s: _batched_face(*s, np.float32), f32))
    direct = rel(_ssprk3(q, u, dt, lambda *s: _batched_face(*s, np.float32, direct=True), f32))
    ref64 = rel(_ssprk3(q, u, dt, lambda *s: _ref_face(*s, np.float64), lambda x: x))
    assert good <= 1e-7, good
    # the direct form loses the limiter's eps scale
    assert direct >= 100.0 * good, (direct, good)
    # the reference's own FP64 evaluation, for scale
    assert ref64 <= 1e-11, ref64
