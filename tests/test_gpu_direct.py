"""The direct FP64 Poisson solve (mode 3) on the GPU through the C-ABI, on
each of its paths: the FFT kernels (each of the three instances, for FFT
lengths up to 1024, 2048 and 4096) with power-of-two transform lengths
(direct_method AUTO) and with other lengths through Bluestein's chirp
z-transform (direct_method FFT), and the dense FP64 products (AUTO on
other lengths, and DENSE).  The potential that poisson_solver solves from
a step's charge is compared with an independent FP64 solve of the same
5-point system (scipy's sparse direct solver), whose right-hand side is
rebuilt exactly from the run's int64 fixed-point charge; on grids where a
sparse factorisation is too slow for a test, the comparison is with a numpy
solve through np.fft and a dense sine basis."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

cp = pytest.importorskip("cupic_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    return cp.init_dev(0)


def solved_run(ncx, ncy, ppc, method=0):
    """A run after init, one push and cc (whose scatter deposits the int64
    charge, read here), then charge_deposition and poisson_solver: the solve
    under test.  Returns the run, the rhs b of the interior rows (q 2^-32 /
    (ds eps0), exact in FP64) and the potential (FP64, with the Dirichlet
    rows), periodic layout."""
    p = cp.params(ncx=ncx, ncy=ncy, ppc=ppc, phi_p=-5.0, jacobi_mode=cp.POISSON_DIRECT,
                  jacobi_tol=1e-11, jacobi_tol_init=1e-11, direct_method=method)
    run = cp.Run(p)
    run.init_sim()
    run.particle_mover()
    run.cc(p.dt)
    q = run.mesh(cp.MESH_CHARGE)[:, :-1].astype(np.float64)
    run.charge_deposition()
    run.poisson_solver()
    b = q * (2.0 ** -32 / (p.ds * p.eps0))
    phi = run.mesh(cp.MESH_PHI)[:, :-1]
    return run, b, phi


def sparse_solve(b, phi, ncx, ncy):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    m = ncy - 1
    idx = np.arange(m * ncx).reshape(m, ncx)
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r.ravel())
        cols.append(c.ravel())
        vals.append(np.full(r.size, v))
    add(idx, idx, 4.0)
    add(idx, np.roll(idx, 1, axis=1), -1.0)
    add(idx, np.roll(idx, -1, axis=1), -1.0)
    add(idx[1:], idx[:-1], -1.0)
    add(idx[:-1], idx[1:], -1.0)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(m * ncx, m * ncx))
    rhs = b[1:ncy].copy()
    rhs[0] += phi[0]
    rhs[m - 1] += phi[ncy]
    return spl.spsolve(A.tocsc(), rhs.ravel()).reshape(m, ncx)


def spectral_solve(b, phi, ncx, ncy):
    """4 phi - neighbours = b: eigenvalues 4 sin^2(pi k / ncx) along x (np.fft)
    plus 4 sin^2(pi (l+1) / 2(m+1)) of the Dirichlet sine basis along y."""
    m = ncy - 1
    rhs = b[1:ncy].copy()
    rhs[0] += phi[0]
    rhs[m - 1] += phi[ncy]
    l = np.arange(m)
    S = np.sqrt(2.0 / (m + 1)) * np.sin(np.pi * np.outer(l + 1, l + 1) / (m + 1))
    F = np.fft.fft(rhs, axis=1)
    k = np.arange(ncx)
    mu = 4 * np.sin(np.pi * k / ncx)[None, :] ** 2 + 4 * np.sin(np.pi * (l + 1) / (2 * (m + 1)))[:, None] ** 2
    G = S @ ((S @ F) / mu)
    return np.fft.ifft(G, axis=1).real


def fft_len(n):
    """The FFT length of an n-point transform: n for a power of two, else
    Bluestein's power of two P >= 2n - 1."""
    if n & (n - 1) == 0:
        return n
    return 1 << (2 * n - 2).bit_length()


def fft_instance(ncx, ncy, method):
    """The FFT kernel instance a grid runs on (the library's fft_class), 0
    for the dense products."""
    pow2 = fft_len(ncx) == ncx and fft_len(2 * ncy) == 2 * ncy
    if method == D or (method == A and not pow2):
        return 0
    n = max(fft_len(ncx), fft_len(2 * ncy))
    return 0 if n > 4096 else (1 if n <= 1024 else (2 if n <= 2048 else 3))


A, D, F = 0, 1, 2  # direct_method: automatic, dense products, FFT for any length
SHAPES = [
    (32, 16, 20, A), (4, 2, 4, A), (4, 8, 8, A), (1024, 4, 2, A), (256, 128, 2, A),  # FFT <= 1024
    (2048, 8, 1, A), (8, 1024, 1, A), (64, 1024, 1, A),                             # FFT <= 2048
    (4096, 2, 1, A),                                                                # FFT <= 4096
    (64, 30, 6, F), (48, 16, 6, F), (37, 9, 7, F), (100, 64, 2, F), (5, 2, 4, F),    # Bluestein
    (600, 200, 1, F), (1000, 3, 1, F),                                              # Bluestein <= 2048
    (2047, 4, 1, F), (6, 1023, 1, F),                                               # Bluestein <= 4096
    (64, 30, 6, A), (48, 16, 6, A), (37, 9, 7, A), (100, 64, 2, A),                 # dense (auto)
    (32, 16, 6, D), (37, 9, 7, D),                                                  # dense (forced)
]
MNAME = {A: "", D: "-dense", F: "-fft"}


@pytest.mark.parametrize("ncx,ncy,ppc,method", SHAPES,
                         ids=[f"{a}x{b}{MNAME[m]}" for a, b, _, m in SHAPES])
def test_direct_solve_matches_sparse_solve(dev, ncx, ncy, ppc, method):
    say(f"{ncx}x{ncy}, method {method}, FFT instance {fft_instance(ncx, ncy, method)}")
    run, b, phi = solved_run(ncx, ncy, ppc, method)
    sweeps, resid = run.last_solve()
    want = sparse_solve(b, phi, ncx, ncy)
    got = phi[1:ncy]
    err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
    say(f"{ncx}x{ncy}: FP64 residual {resid:.3e}, relative difference to the sparse solve {err:.3e}")
    assert resid < 1e-11
    assert err < 1e-11
    assert np.all(np.isfinite(phi))
    run.close()


@pytest.mark.parametrize("ncx,ncy", [(1000, 500), (1023, 511)])
def test_direct_solve_bluestein_serving_grids(dev, ncx, ncy):
    """Serving-sized grids that are not powers of two (FFT length 2048 both
    ways) against the numpy spectral solve."""
    run, b, phi = solved_run(ncx, ncy, 1, F)
    sweeps, resid = run.last_solve()
    want = spectral_solve(b, phi, ncx, ncy)
    got = phi[1:ncy]
    err = np.abs(got - want).max() / np.abs(want).max()
    say(f"{ncx}x{ncy}: FP64 residual {resid:.3e}, relative difference to numpy {err:.3e}")
    assert resid < 1e-11
    assert err < 1e-11
    run.close()


def test_direct_solve_benchmark_grid(dev):
    ncx, ncy = 1024, 512
    run, b, phi = solved_run(ncx, ncy, 2)
    sweeps, resid = run.last_solve()
    want = spectral_solve(b, phi, ncx, ncy)
    got = phi[1:ncy]
    err = np.abs(got - want).max() / np.abs(want).max()
    say(f"{ncx}x{ncy}: FP64 residual {resid:.3e}, relative difference to numpy {err:.3e}")
    assert resid < 1e-11
    assert err < 1e-11
    run.close()


def test_direct_solve_batch_mixes_paths(dev):
    """A batch holding runs of every FFT instance, power-of-two and
    Bluestein lengths, and of the dense products (automatic and forced)
    steps each of them bit for bit as the run stepped alone."""
    shapes = [dict(ncx=64, ncy=32, ppc=6), dict(ncx=2048, ncy=8, ppc=2, seed=5),
              dict(ncx=40, ncy=20, ppc=6, seed=7), dict(ncx=16, ncy=1024, ppc=1, seed=9),
              dict(ncx=40, ncy=20, ppc=6, seed=11, direct_method=F),
              dict(ncx=4096, ncy=3, ppc=1, seed=13, direct_method=F),
              dict(ncx=99, ncy=45, ppc=5, seed=15, direct_method=F),
              dict(ncx=64, ncy=32, ppc=6, seed=17, direct_method=D)]
    alone, mixed = [], []
    for kw in shapes:
        p = cp.params(**kw, phi_p=-5.0, jacobi_mode=cp.POISSON_DIRECT, jacobi_tol=1e-11,
                      jacobi_tol_init=1e-11)
        for lst in (alone, mixed):
            r = cp.Run(p)
            r.init_sim()
            lst.append(r)
    for r in alone:
        t = r.p.dt
        for _ in range(4):
            r.step(t)
            t += r.p.dt
    b = cp.Batch(mixed, graph=True)
    b.step(4)
    b.close()
    for ra, rm, kw in zip(alone, mixed, shapes):
        assert np.array_equal(ra.mesh(cp.MESH_PHI), rm.mesh(cp.MESH_PHI)), kw
        for s in (0, 1):
            for u, v in zip(ra.particles(s), rm.particles(s)):
                assert np.array_equal(u, v), kw
        say(f"{kw['ncx']}x{kw['ncy']}: batched == alone")
    for r in alone + mixed:
        r.close()
