"""Pins for the oracle's RNG, parameter derivation and injection clock.

Philox-4x64-10 is checked against numpy.random.Philox (a library
implementation of the same published generator) and the Random123
known-answer vector; derived parameters against the closed forms of
src/init.cu:136, 686-714.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

L = O.lib()


def philox(ctr, key):
    out = np.zeros(4, np.uint64)
    L.orc_philox4x64_10(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64), out)
    return out


def test_philox_random123_known_answer():
    # Random123 kat_vectors: philox4x64_10, counter 0, key 0
    got = philox([0, 0, 0, 0], [0, 0])
    want = [0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    assert [int(v) for v in got] == want


@pytest.mark.parametrize("key", [0, 1, 20261203, (7 << 64) | 123456789, 2**128 - 1])
@pytest.mark.parametrize("ctr0", [0, 5, 2**40 + 17])
def test_philox_matches_numpy(key, ctr0):
    # numpy increments the 256-bit counter before producing each block
    bg = np.random.Philox(key=key, counter=ctr0)
    raw = bg.random_raw(12)
    k = [key & (2**64 - 1), key >> 64]
    for b in range(3):
        c = ctr0 + 1 + b
        ctr = [c & (2**64 - 1), (c >> 64) & (2**64 - 1), 0, 0]
        assert [int(v) for v in philox(ctr, k)] == [int(v) for v in raw[4 * b:4 * b + 4]]


def test_u01_open_interval_and_exactness():
    assert L.orc_u01_open(0) == 0.5 * 2.0**-52
    assert L.orc_u01_open(2**64 - 1) == 1.0 - 0.5 * 2.0**-52
    assert 0.0 < L.orc_u01_open(123) < 1.0


def test_params_closed_forms():
    p = O.make_params()  # config 1
    assert p.Lx == 64.0 and p.Ly == 32.0
    assert p.per_row == 2560 and p.per_row * p.ncy == 163840
    assert p.eps0 == p.n == 160.0
    assert p.dtin_e == pytest.approx(math.sqrt(2 * math.pi) / (160 * 64 * 0.5), rel=1e-15)
    assert p.dtin_i == pytest.approx(p.dtin_e * math.sqrt(100.0 / 1.0), rel=1e-14)
    # ~102.1 electrons and ~10.2 ions per step (one-sided Maxwellian flux)
    assert p.dt / p.dtin_e == pytest.approx(102.13, abs=0.01)
    assert p.dt / p.dtin_i == pytest.approx(10.213, abs=0.001)
    p2 = O.make_params(ncx=1024, ncy=512, ppc=200, mi=1836.0, kti=0.1, phi_p=-10.0)
    assert p2.per_row * p2.ncy == 104857600


def test_injection_count_recurrence():
    """in = 1 + int((t+dt-tin)/dtin) with tin advanced once per particle
    (src/cc.cu:179, 247): over K steps the total equals the number of
    insertion times tin0 + j*dtin below the last t+dt."""
    import ctypes as C
    dt, dtin = 0.05, 4.895758348888673e-4
    t = dt
    tin = C.c_double(t + dtin)
    tin0 = tin.value
    total = 0
    seq = np.zeros(4096)
    for k in range(2000):
        n = L.orc_injection_count(t, dt, C.byref(tin), dtin, seq.ctypes.data, 4096)
        assert 101 <= n <= 104 or k == 0
        # every insertion time of this step lies in [.., t+dt)
        assert np.all(seq[:n] < t + dt)
        total += n
        t += dt
    expect = math.floor((t - tin0) / dtin) + 1  # t is now last t + dt
    assert abs(total - expect) <= 1
    assert tin.value >= t - 1e-9


def test_initial_load_distributions():
    p = O.make_params(ncx=64, ncy=32, ppc=20)
    N = p.per_row * p.ncy
    a = [np.empty(N) for _ in range(4)]
    L.orc_create_particles_f64(p, 0, *a)
    x, y, vx, vy = a
    # exact per-row counts (src/init.cu:159-164)
    rows = np.floor(y / p.ds).astype(int)
    assert np.array_equal(np.bincount(rows, minlength=p.ncy), np.full(p.ncy, p.per_row))
    assert stats.kstest(x / p.Lx, "uniform").pvalue > 1e-3
    assert stats.kstest(vx, "norm").pvalue > 1e-3
    assert stats.kstest(vy, "norm").pvalue > 1e-3
    b = [np.empty(N) for _ in range(4)]
    L.orc_create_particles_f64(p, 1, *b)
    sig_i = math.sqrt(p.kti / p.mi)
    assert stats.kstest(b[2] / sig_i, "norm").pvalue > 1e-3
    # the two species use independent streams
    assert not np.array_equal(x, b[0])


def test_injected_velocity_distributions():
    p = O.make_params(ncx=64, ncy=32)
    n = 20000
    tin = np.linspace(0.05, 0.1, n, endpoint=False)
    nodes = O.nodes(p)
    Ex = np.zeros(nodes)
    Ey = np.zeros(nodes)
    out = [np.empty(n) for _ in range(4)]
    L.orc_inject_f64(p, 0, 0, n, 0.05, tin, Ex, Ey, *out)
    x, y, vx, vy = out
    assert stats.kstest(vx, "norm").pvalue > 1e-3                      # N(0, 1)
    assert stats.kstest(-vy, "rayleigh").pvalue > 1e-3                 # -Rayleigh(1)
    # with E = 0 the partial push is free flight over t+dt-tin
    fpt = 0.05 + p.dt
    assert np.allclose(y, p.Ly + (fpt - tin) * vy, rtol=0, atol=1e-12)
    assert np.all((x >= 0) & (x < p.Lx))


def test_injected_fp32_positions_stay_inside():
    """Rounded to FP32, an injected y just below Ly can reach Ly (256 here, whose
    ulp is 2^-15): it must become the largest float below Ly, never Ly."""
    p = O.make_params(ncx=64, ncy=512)
    n = 200000
    t = 0.05
    tin = np.linspace(t + p.dt, t, n, endpoint=False)   # partial push down to 0
    E = np.zeros(2 * O.nodes(p), np.float32)
    out = [np.empty(n, np.float32) for _ in range(4)]
    L.orc_inject_f32(p, 0, 0, n, t, tin, E, *out)
    x, y, vx, vy = out
    Ly = np.float32(p.Ly)
    below = np.nextafter(Ly, np.float32(0))
    assert np.all((x >= 0) & (x < np.float32(p.Lx)))
    assert np.all((y >= 0) & (y < Ly))
    assert np.count_nonzero(y == below) > 0       # the clamp was exercised
    # the FP64 positions round to the same floats wherever they stay below Ly
    y64 = np.empty(n)
    o64 = [np.empty(n) for _ in range(4)]
    nodes = O.nodes(p)
    L.orc_inject_f64(p, 0, 0, n, t, tin, np.zeros(nodes), np.zeros(nodes), *o64)
    y64 = o64[1]
    keep = y64.astype(np.float32) < Ly
    assert np.array_equal(y[keep], y64[keep].astype(np.float32))
