"""Pins for the two velocity kicks outside the leapfrog step.

* The initial half-step stagger, v(-dt/2) = v(0) - (dt/2) q E / m
  (src/init.cu:825-826), in its FP32 mirror form ``orc_mirror_half_kick``:
  exact against the closed form on dyadic inputs, and within FP32 rounding
  of the FP64 ``orc_half_kick_f64`` on random states and fields.
* The partial push of injected particles (src/cc.cu:222-244): the position
  moves with the drawn velocity over fpt - tin = t + dt - tin, and the
  velocity takes q E / m over fvt - tin = t + dt/2 - tin, with E
  interpolated on the top node row (disty = 1).  Under the Q7 quirk flag the
  kick is subtracted for both species, as the committed code does.

Each closed form is evaluated here from the mathematics, not from the
oracle's code path, so a wrong sign, a wrong time factor (fpt for fvt, a
missing 1/2), a wrong mass or the wrong node row fails a test.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()


def interleave(Ex, Ey):
    E = np.empty(2 * Ex.size, np.float32)
    E[0::2] = Ex.ravel()
    E[1::2] = Ey.ravel()
    return E


# ---- initial half kick (src/init.cu:825-826) -------------------------------

@pytest.mark.parametrize("species", [0, 1])
def test_mirror_half_kick_dyadic_closed_form(species):
    """dt = 2^-4 and mi = 4 make q dt / (2 m) a power of two; E, v and the
    in-cell fractions are dyadic, so every FP32 operation is exact and the
    result must equal v - (dt/2) q E / m exactly."""
    p = O.make_params(ncx=16, ncy=8, dt=0.0625, mi=4.0)
    c = O.f32c(p)
    q = -1.0 if species == 0 else 1.0
    m = 1.0 if species == 0 else 4.0
    ex0, ey0 = 0.375, -1.25
    nn = O.nodes(p)
    E = interleave(np.full(nn, ex0, np.float32), np.full(nn, ey0, np.float32))
    r = np.random.default_rng(11)
    n = 512
    # positions on quarter-cell fractions: CIC weights are exact
    x = ((r.integers(0, p.ncx * 4, n) + 0.0) * p.ds / 4).astype(np.float32)
    y = ((r.integers(0, p.ncy * 4, n) + 0.0) * p.ds / 4).astype(np.float32)
    vx = (r.integers(-64, 64, n) / 32.0).astype(np.float32)
    vy = (r.integers(-64, 64, n) / 32.0).astype(np.float32)
    wx = vx.astype(np.float64) - 0.5 * p.dt * q * ex0 / m
    wy = vy.astype(np.float64) - 0.5 * p.dt * q * ey0 / m
    L.orc_mirror_half_kick(c, species, n, x, y, vx, vy, E)
    assert np.array_equal(vx.astype(np.float64), wx)
    assert np.array_equal(vy.astype(np.float64), wy)


@pytest.mark.parametrize("species", [0, 1])
def test_mirror_half_kick_matches_fp64(species):
    """Random states and a random field: the FP32 mirror stays within FP32
    rounding of the FP64 half kick (pinned above and by its own closed form)."""
    p = O.make_params(ncx=32, ncy=16, mi=100.0)
    c = O.f32c(p)
    q, m = (-1.0, 1.0) if species == 0 else (1.0, 100.0)
    nn = O.nodes(p)
    r = np.random.default_rng(5 + species)
    Ex = r.normal(size=nn).astype(np.float32)
    Ey = r.normal(size=nn).astype(np.float32)
    n = 4000
    x = (r.random(n) * p.Lx).astype(np.float32)
    y = (r.random(n) * p.Ly).astype(np.float32)
    vx = r.normal(size=n).astype(np.float32)
    vy = r.normal(size=n).astype(np.float32)
    d = [a.astype(np.float64) for a in (x, y, vx, vy)]
    v0 = d[2].copy(), d[3].copy()
    L.orc_half_kick_f64(n, d[0], d[1], d[2], d[3], Ex.astype(np.float64).copy(),
                        Ey.astype(np.float64).copy(), p.ncx, p.ncy, p.ds, q, m, p.dt)
    L.orc_mirror_half_kick(c, species, n, x, y, vx, vy, interleave(Ex, Ey))
    kick = 0.5 * p.dt / m * 4.0          # bound on |dt q E / (2m)| for |E| < 4
    for got, want, start in ((vx, d[2], v0[0]), (vy, d[3], v0[1])):
        assert np.abs(want - start).max() > 0.2 * kick      # the kick is not zero
        # FP32: the start value's rounding plus a few ulps of the kick
        tol = 4 * np.spacing(np.abs(want).astype(np.float32)).astype(np.float64) + 1e-6 * kick
        assert np.all(np.abs(got.astype(np.float64) - want) <= tol)


# ---- injected particles (src/cc.cu:214-247) ----------------------------------

def inject64(p, species, t, tin, Ex, Ey):
    n = len(tin)
    out = [np.empty(n) for _ in range(4)]
    L.orc_inject_f64(p, species, 0, n, t, np.ascontiguousarray(tin), Ex, Ey, *out)
    return out


def kick_sign(species, quirks):
    """+q for the physical kick; -1 for both species under the Q7 quirk."""
    if quirks & 1:          # ORC_QUIRK_Q7_REFERENCE_KICK
        return -1.0
    return -1.0 if species == 0 else 1.0


@pytest.mark.parametrize("quirks", [0, 1])
@pytest.mark.parametrize("species", [0, 1])
def test_injection_kick_uniform_field(species, quirks):
    """Uniform E: positions are those of free flight over fpt - tin with the
    drawn velocity, and the velocity changes by exactly
    sign * (t + dt/2 - tin) * E / m."""
    p = O.make_params(ncx=64, ncy=32, mi=100.0, kti=0.5, quirks=quirks)
    m = 1.0 if species == 0 else p.mi
    t = 0.35
    n = 3000
    tin = t + p.dt * np.linspace(0.0, 1.0, n, endpoint=False)[::-1]
    nn = O.nodes(p)
    z = np.zeros(nn)
    x0, y0, vx0, vy0 = inject64(p, species, t, tin, z, z)
    # free flight: y = Ly + (t + dt - tin) * vy
    assert np.allclose(y0, p.Ly + (t + p.dt - tin) * vy0, rtol=0, atol=1e-12)
    ex0, ey0 = 0.7, -2.3
    x, y, vx, vy = inject64(p, species, t, tin, np.full(nn, ex0), np.full(nn, ey0))
    assert np.array_equal(x, x0) and np.array_equal(y, y0)   # E does not move positions
    sgn = kick_sign(species, quirks)
    fvt = t + 0.5 * p.dt
    assert np.allclose(vx - vx0, sgn * (fvt - tin) * ex0 / m, rtol=1e-9, atol=1e-14)
    assert np.allclose(vy - vy0, sgn * (fvt - tin) * ey0 / m, rtol=1e-9, atol=1e-14)
    # the kick interval spans both signs of fvt - tin: early and late insertions
    assert (fvt - tin).min() < 0 < (fvt - tin).max()


@pytest.mark.parametrize("quirks", [0, 1])
@pytest.mark.parametrize("species", [0, 1])
def test_injection_kick_linear_field_on_top_row(species, quirks):
    """E linear in x along the top node row j = ncy and large, unrelated
    values on row ncy - 1: the kick must use the top row only (disty = 1)
    and interpolate it linearly at the drawn x, E(x) = a + b x / ds."""
    p = O.make_params(ncx=64, ncy=32, mi=100.0, kti=0.5, quirks=quirks)
    m = 1.0 if species == 0 else p.mi
    t = 1.15
    n = 4000
    tin = t + p.dt * np.linspace(0.0, 1.0, n, endpoint=False)
    nn = O.nodes(p)
    z = np.zeros(nn)
    x0, _, vx0, vy0 = inject64(p, species, t, tin, z, z)
    i = np.arange(p.ncx)
    ax, bx, ay, by = 0.4, 0.03, -1.1, -0.02
    Ex = np.full((p.ncy + 1, p.ncx), 1e3)
    Ey = np.full((p.ncy + 1, p.ncx), -1e3)
    Ex[p.ncy] = ax + bx * i
    Ey[p.ncy] = ay + by * i
    x, _, vx, vy = inject64(p, species, t, tin, Ex.ravel().copy(), Ey.ravel().copy())
    # the drawn x: undo the free flight over fpt - tin, modulo Lx
    xd = np.mod(x0 - (t + p.dt - tin) * vx0, p.Lx)
    inner = xd < (p.ncx - 1) * p.ds - 1e-9      # the last cell wraps to node 0
    assert inner.sum() > 0.9 * n
    sgn = kick_sign(species, quirks)
    w = sgn * (t + 0.5 * p.dt - tin) / m
    ex = ax + bx * xd / p.ds
    ey = ay + by * xd / p.ds
    assert np.allclose((vx - vx0)[inner], (w * ex)[inner], rtol=1e-8, atol=1e-12)
    assert np.allclose((vy - vy0)[inner], (w * ey)[inner], rtol=1e-8, atol=1e-12)
    # in the last cell the top row interpolates between node ncx-1 and node 0
    last = ~inner & (xd < p.Lx)
    f = xd[last] / p.ds - (p.ncx - 1)
    ex_last = (ax + bx * (p.ncx - 1)) * (1 - f) + ax * f
    assert np.allclose((vx - vx0)[last], w[last] * ex_last, rtol=1e-8, atol=1e-12)


def test_injection_fp32_is_the_rounded_fp64_state():
    """The device-order FP32 injection is the FP64 state rounded once, with a
    field given in FP32."""
    p = O.make_params(ncx=64, ncy=32, mi=100.0)
    t = 0.5
    n = 2000
    tin = t + p.dt * np.linspace(0.0, 1.0, n, endpoint=False)
    r = np.random.default_rng(2)
    nn = O.nodes(p)
    Ex = r.normal(size=nn).astype(np.float32)
    Ey = r.normal(size=nn).astype(np.float32)
    for s in (0, 1):
        o64 = inject64(p, s, t, tin, Ex.astype(np.float64), Ey.astype(np.float64))
        o32 = [np.empty(n, np.float32) for _ in range(4)]
        L.orc_inject_f32(p, s, 0, n, t, tin, interleave(Ex, Ey), *o32)
        keep = o64[1].astype(np.float32) < np.float32(p.Ly)
        for a, b in zip(o32, o64):
            assert np.array_equal(a[keep], b[keep].astype(np.float32))
