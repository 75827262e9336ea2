"""Pins for the oracle's mesh layer: deposition (FP64 and int64 fixed point),
Jacobi Poisson solve, E = -grad(phi).

Pins come from outside the oracle: exact conservation laws, closed-form
single-particle weights, the discrete Laplacian applied by numpy, scipy's
sparse direct solver, and closed-form finite differences.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O

L = O.lib()


def small(ncx=16, ncy=8, **kw):
    return O.make_params(ncx=ncx, ncy=ncy, **kw)


def dep64(p, x, y, q):
    Q = np.zeros(O.nodes(p))
    L.orc_deposit_f64(len(x), np.ascontiguousarray(x, np.float64),
                      np.ascontiguousarray(y, np.float64), p.ds, p.ncx, p.ncy, q, Q)
    return Q.reshape(p.ncy + 1, p.ncx)


def depfix(p, x, y, sign):
    c = O.f32c(p)
    acc = np.zeros(O.nodes(p), np.int64)
    L.orc_deposit_fixed_f32(len(x), np.ascontiguousarray(x, np.float32),
                            np.ascontiguousarray(y, np.float32), c, sign, acc)
    return acc.reshape(p.ncy + 1, p.ncx)


def random_positions(p, n, seed=1):
    r = np.random.default_rng(seed)
    return (r.random(n) * p.Lx).astype(np.float32), (r.random(n) * p.Ly).astype(np.float32)


# ---------------------------------------------------------------- deposition

def test_fixed_point_conserves_charge_exactly():
    p = small()
    x, y = random_positions(p, 5000)
    acc = depfix(p, x, y, -1)
    assert acc.sum() == -(2**32) * 5000
    assert depfix(p, x, y, +1).sum() == (2**32) * 5000


def test_fixed_point_single_particle_weights():
    p = small()
    # cell (3, 2) origin at (1.5, 1.0); fractions 0.25 and 0.75 are exact
    x = np.array([1.5 + 0.25 * p.ds], np.float32)
    y = np.array([1.0 + 0.75 * p.ds], np.float32)
    acc = depfix(p, x, y, 1)
    gx, wx, gy, wy = 49152, 16384, 16384, 49152
    want = np.zeros_like(acc)
    want[2, 3], want[2, 4] = gx * gy, wx * gy
    want[3, 3], want[3, 4] = gx * wy, wx * wy
    assert np.array_equal(acc, want)


def test_fixed_point_periodic_fold():
    p = small()
    # a particle in the last column puts its right-hand charge on column 0
    x = np.array([p.Lx - 0.5 * p.ds], np.float32)
    y = np.array([0.5 * p.ds], np.float32)
    acc = depfix(p, x, y, 1)
    assert acc[0, p.ncx - 1] == 2**30 and acc[0, 0] == 2**30
    assert acc[1, p.ncx - 1] == 2**30 and acc[1, 0] == 2**30


def test_fixed_point_matches_fp64_within_quantum():
    p = small()
    x, y = random_positions(p, 3000, seed=3)
    acc = depfix(p, x, y, 1)
    Q = dep64(p, x.astype(np.float64), y.astype(np.float64), 1.0)
    # each vertex weight is quantised to 2^-16 per axis: |error| <= 2^-16 per particle
    err = np.abs(acc / 2.0**32 - Q)
    assert err.max() <= 3000 * 2.0**-16
    assert abs(Q.sum() - 3000) < 1e-9


def test_fp64_deposit_lattice_uniform():
    """Particles at every cell centre give each interior node exactly 1."""
    p = small()
    ii, jj = np.meshgrid(np.arange(p.ncx), np.arange(p.ncy))
    x = ((ii + 0.5) * p.ds).ravel()
    y = ((jj + 0.5) * p.ds).ravel()
    Q = dep64(p, x, y, 1.0)
    assert np.allclose(Q[1:-1], 1.0, rtol=0, atol=1e-15)
    assert np.allclose(Q[0], 0.5) and np.allclose(Q[-1], 0.5)
    rho = np.zeros(O.nodes(p))
    L.orc_rho_from_Q(p.ncx, p.ncy, p.ds, Q.ravel().copy(), rho)
    rho = rho.reshape(Q.shape)
    v = p.ds**3
    assert np.allclose(rho[1:-1], 1.0 / v) and np.allclose(rho[0], 1.0 / v)


def test_fp64_deposit_brute_force_tiny():
    p = small(ncx=4, ncy=3)
    r = np.random.default_rng(7)
    x = r.random(50) * p.Lx
    y = r.random(50) * p.Ly
    Q = dep64(p, x, y, -1.0)
    want = np.zeros((p.ncy + 1, p.ncx))
    for xi, yi in zip(x, y):
        # bilinear hat functions evaluated directly at every node
        for j in range(p.ncy + 1):
            for i in range(p.ncx + 1):
                wxh = max(0.0, 1 - abs(xi - i * p.ds) / p.ds)
                wyh = max(0.0, 1 - abs(yi - j * p.ds) / p.ds)
                want[j, i % p.ncx] += -wxh * wyh
    assert np.allclose(Q, want, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- Jacobi

def laplacian_matrix(ncx, ncy):
    """5-point operator on interior rows 1..ncy-1, periodic in x, acting on
    the unknowns; Dirichlet rows enter the right-hand side."""
    n = ncx * (ncy - 1)
    idx = lambda i, j: (j - 1) * ncx + (i % ncx)
    rows, cols, vals = [], [], []
    for j in range(1, ncy):
        for i in range(ncx):
            k = idx(i, j)
            rows += [k]; cols += [k]; vals += [4.0]
            for (a, b) in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                if 1 <= b <= ncy - 1:
                    rows.append(k); cols.append(idx(a, b)); vals.append(-1.0)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def solve_direct(p, rho, phi_bc):
    A = laplacian_matrix(p.ncx, p.ncy)
    b = (p.ds**2 * rho[1:-1] / p.eps0).ravel().copy()
    b[:p.ncx] += phi_bc[0]
    b[-p.ncx:] += phi_bc[-1]
    sol = spla.spsolve(A.tocsc(), b)
    out = phi_bc.copy()
    out[1:-1] = sol.reshape(p.ncy - 1, p.ncx)
    return out


def test_linear_phi_is_exact_fixed_point():
    p = small(phi_p=-5.0)
    phi = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi)
    rho = np.zeros(O.nodes(p))
    assert L.orc_residual_max(p.ncx, p.ncy, p.ds, p.eps0, rho, phi) < 1e-15
    phi2 = phi.copy()
    work = np.zeros_like(phi)
    L.orc_jacobi_sweeps(p.ncx, p.ncy, p.ds, p.eps0, rho, phi2, work, 10, None)
    assert np.abs(phi2 - phi).max() < 1e-14
    # and it matches the closed form of src/init.cu:239
    j = np.arange(p.ncy + 1)
    assert np.allclose(phi.reshape(p.ncy + 1, p.ncx), ((1 - j / p.ncy) * p.phi_p)[:, None])


def test_manufactured_solution_and_scipy_direct_solve():
    p = small(ncx=12, ncy=6, phi_p=-5.0)
    sh = (p.ncy + 1, p.ncx)
    jj, ii = np.mgrid[0:p.ncy + 1, 0:p.ncx]
    star = np.sin(2 * np.pi * ii / p.ncx) * np.sin(np.pi * jj / p.ncy) \
        + (1 - jj / p.ncy) * p.phi_p
    # rho from the discrete Laplacian of star, applied here with numpy rolls
    lap = (np.roll(star, 1, 1) + np.roll(star, -1, 1) - 4 * star)
    lap[1:-1] += star[:-2] + star[2:]
    rho = np.zeros(sh)
    rho[1:-1] = -lap[1:-1] * p.eps0 / p.ds**2
    direct = solve_direct(p, rho, star)
    assert np.abs(direct - star).max() < 1e-12
    phi = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi)
    import ctypes as C
    res = C.c_double()
    s = L.orc_poisson_solve(p.ncx, p.ncy, p.ds, p.eps0, rho.ravel().copy(), phi, 1, 1e-12,
                            0, 200000, 50, C.byref(res))
    assert res.value <= 1e-12 and 0 < s < 200000
    assert np.abs(phi.reshape(sh) - star).max() < 1e-9


@pytest.mark.parametrize("ncx,ncy", [(16, 8), (12, 6), (37, 9), (64, 2), (5, 4), (128, 64)])
def test_direct_mode_matches_sparse_solve(ncx, ncy):
    """mode 3 (FFT or direct DFT along x, Thomas along y) against scipy's
    sparse factorisation of the same 5-point system, and against the
    manufactured solution, for even, odd, power-of-two and other ncx."""
    p = small(ncx=ncx, ncy=ncy, phi_p=-5.0)
    sh = (p.ncy + 1, p.ncx)
    r = np.random.default_rng(ncx * 100 + ncy)
    rho = np.zeros(sh)
    rho[1:-1] = r.normal(size=(p.ncy - 1, p.ncx)) * 5
    phi = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi)
    phi.reshape(sh)[0] += r.normal(size=p.ncx)   # non-uniform Dirichlet rows
    phi.reshape(sh)[-1] += r.normal(size=p.ncx)
    want = solve_direct(p, rho, phi.reshape(sh).copy())
    import ctypes as C
    res = C.c_double()
    s = L.orc_poisson_solve(p.ncx, p.ncy, p.ds, p.eps0, rho.ravel().copy(), phi, 3, 0.0, 0, 0,
                            50, C.byref(res))
    assert s == 0
    got = phi.reshape(sh)
    assert np.abs(got - want).max() < 1e-10 * max(1.0, np.abs(want).max())
    assert res.value < 1e-10
    # the manufactured solution of test_manufactured_solution_and_scipy_direct_solve
    jj, ii = np.mgrid[0:p.ncy + 1, 0:p.ncx]
    star = np.sin(2 * np.pi * ii / p.ncx) * np.sin(np.pi * jj / p.ncy) \
        + (1 - jj / p.ncy) * p.phi_p
    lap = (np.roll(star, 1, 1) + np.roll(star, -1, 1) - 4 * star)
    lap[1:-1] += star[:-2] + star[2:]
    rho2 = np.zeros(sh)
    rho2[1:-1] = -lap[1:-1] * p.eps0 / p.ds**2
    phi2 = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi2)
    L.orc_poisson_solve(p.ncx, p.ncy, p.ds, p.eps0, rho2.ravel().copy(), phi2, 3, 0.0, 0, 0, 50,
                        C.byref(res))
    assert np.abs(phi2.reshape(sh) - star).max() < 1e-11


def test_reference_stopping_rule_min_sweeps():
    """mode 0: at least min_sweeps sweeps, then until max|dphi| < tol
    (src/mesh.cu:66-67, 94-113)."""
    p = small(phi_p=-5.0)
    phi = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi)
    rho = np.zeros(O.nodes(p))
    import ctypes as C
    res = C.c_double()
    s = L.orc_poisson_solve(p.ncx, p.ncy, p.ds, p.eps0, rho, phi, 0, 1e-4, 18, 0, 50,
                            C.byref(res))
    assert s == 18  # already converged: exactly the minimum
    r = np.random.default_rng(2)
    rho = r.normal(size=O.nodes(p)) * 10
    phi0 = phi.copy()
    s = L.orc_poisson_solve(p.ncx, p.ncy, p.ds, p.eps0, rho, phi, 0, 1e-4, 18, 0, 50,
                            C.byref(res))
    assert s > 18 and res.value < 1e-4
    # the last sweep's update is below tol, the one before it is not
    phi1 = phi0.copy()
    work = np.zeros_like(phi1)
    last = C.c_double()
    L.orc_jacobi_sweeps(p.ncx, p.ncy, p.ds, p.eps0, rho, phi1, work, s - 1, C.byref(last))
    assert last.value >= 1e-4
    assert np.isfinite(phi1).all()


def test_jacobi_spectral_radius_matches_theory():
    """Error in the x-uniform lowest y mode decays by (1+cos(pi/ncy))/2
    per sweep -- the slowest Jacobi mode with periodic x."""
    p = small(ncx=8, ncy=16, phi_p=0.0)
    sh = (p.ncy + 1, p.ncx)
    jj = np.mgrid[0:p.ncy + 1, 0:p.ncx][0]
    phi = (np.sin(np.pi * jj / p.ncy)).ravel().copy()
    rho = np.zeros(O.nodes(p))
    work = np.zeros_like(phi)
    L.orc_jacobi_sweeps(p.ncx, p.ncy, p.ds, p.eps0, rho, phi, work, 40, None)
    lam = 0.5 * (1 + math.cos(math.pi / p.ncy))
    expect = lam**40 * np.sin(np.pi * jj / p.ncy)
    assert np.abs(phi.reshape(sh) - expect).max() < 1e-12


# ---------------------------------------------------------------- field

def field(p, phi):
    Ex = np.zeros(O.nodes(p))
    Ey = np.zeros(O.nodes(p))
    L.orc_field(p.ncx, p.ncy, p.ds, np.ascontiguousarray(phi).ravel().copy(), Ex, Ey)
    sh = (p.ncy + 1, p.ncx)
    return Ex.reshape(sh), Ey.reshape(sh)


def test_field_linear_phi_exact_everywhere():
    p = small(phi_p=-5.0)
    phi = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi)
    Ex, Ey = field(p, phi)
    assert np.abs(Ex).max() == 0.0
    assert np.allclose(Ey, p.phi_p / p.Ly, rtol=1e-13, atol=0)  # one-sided rows included


def test_field_sinusoid_closed_form():
    p = small(ncx=32, ncy=4)
    ii = np.mgrid[0:p.ncy + 1, 0:p.ncx][1]
    k = 2 * np.pi / p.ncx
    phi = np.sin(k * ii)
    Ex, Ey = field(p, phi)
    want = -np.cos(k * ii) * np.sin(k) / p.ds
    assert np.allclose(Ex[1:-1], want[1:-1], rtol=0, atol=1e-13)
    assert np.all(Ex[0] == 0) and np.all(Ex[-1] == 0)
    assert np.abs(Ey).max() < 1e-13


# ---------------------------------------------------------------- FP32 mirror solve

def mirror_case(ncx=24, ncy=12, n=3000):
    import ctypes as C
    p = small(ncx=ncx, ncy=ncy)
    x, y = random_positions(p, n, seed=5)
    acc = depfix(p, x, y, 1).ravel()          # ions only: a net charge
    b64 = np.empty(O.nodes(p))
    b32 = np.empty(O.nodes(p), np.float32)
    L.orc_mirror_rhs(p, acc, b64, b32)
    phi = np.zeros(O.nodes(p))
    L.orc_phi_linear(p.ncx, p.ncy, p.phi_p, phi)
    return C, p, acc, b64, b32, phi


def numpy_residual(p, b64, phi32):
    ph = phi32.astype(np.float64).reshape(p.ncy + 1, p.ncx)
    b = b64.reshape(p.ncy + 1, p.ncx)
    s = (((b + np.roll(ph, 1, 1)) + np.roll(ph, -1, 1)) + np.roll(ph, -1, 0)) + np.roll(ph, 1, 0)
    return float(np.abs(0.25 * (s - 4.0 * ph))[1:p.ncy].max())


def test_mirror_rhs_is_charge_over_ds_eps0():
    C, p, acc, b64, b32, _ = mirror_case()
    want = acc.astype(np.float64) * 2.0 ** -32 / (p.ds * p.n)
    want = want.reshape(p.ncy + 1, p.ncx)
    want[0] = want[-1] = 0.0
    assert np.allclose(b64.reshape(want.shape), want, rtol=1e-15, atol=0)
    assert np.array_equal(b32, b64.astype(np.float32))


def test_mirror_solve_matches_fp64_sweeps():
    """The FP32 sweeps track the FP64 Jacobi iterates within FP32 rounding."""
    C, p, acc, b64, b32, phi = mirror_case()
    phi32 = phi.astype(np.float32)
    res = C.c_double()
    sw = L.orc_mirror_solve_f32(p.ncx, p.ncy, b64, b32, phi32, 0.0, 200, 50, C.byref(res))
    assert sw == 200
    rho = np.zeros(O.nodes(p))
    L.orc_rho_from_Q(p.ncx, p.ncy, p.ds, acc.astype(np.float64) * 2.0 ** -32, rho)
    work = np.zeros_like(phi)
    last = C.c_double()
    L.orc_jacobi_sweeps(p.ncx, p.ncy, p.ds, p.eps0, rho, phi, work, 200, C.byref(last))
    assert np.abs(phi32 - phi).max() / np.abs(phi).max() < 2e-6
    assert res.value == numpy_residual(p, b64, phi32)


def test_mirror_solve_stopping_rule():
    C, p, acc, b64, b32, phi = mirror_case()
    r0 = numpy_residual(p, b64, phi.astype(np.float32))
    res = C.c_double()
    phi32 = phi.astype(np.float32)
    assert L.orc_mirror_solve_f32(p.ncx, p.ncy, b64, b32, phi32, r0, 500, 50, C.byref(res)) == 0
    assert res.value == r0
    # a tolerance reached mid-way stops at the first check point below it
    phi32 = phi.astype(np.float32)
    full = []
    for k in range(1, 7):
        q = phi.astype(np.float32)
        L.orc_mirror_solve_f32(p.ncx, p.ncy, b64, b32, q, 0.0, 50 * k, 50, C.byref(res))
        full.append(res.value)
    tol = 0.5 * (full[2] + full[3])        # between the checks at 150 and 200 sweeps
    assert full[3] < tol < full[2]
    sw = L.orc_mirror_solve_f32(p.ncx, p.ncy, b64, b32, phi32, tol, 500, 50, C.byref(res))
    assert sw == 200 and res.value == full[3]
