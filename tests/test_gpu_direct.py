"""The direct FP64 Poisson solve (mode 3) on the GPU through the C-ABI, on
each of its paths: the FFT kernels for ncx and 2 ncy powers of two (the
instance for transform lengths up to 1024 and the one up to 4096) and the
dense products for other shapes.  The potential that poisson_solver solves
from a step's charge is compared with an independent FP64 solve of the
same 5-point system (scipy's sparse direct solver), whose right-hand side is
rebuilt exactly from the run's int64 fixed-point charge; at the benchmark's
1024 x 512 cells, where a sparse factorisation is too slow for a test, the
comparison is with a numpy solve through np.fft and a dense sine basis."""
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


def solved_run(ncx, ncy, ppc):
    """A run after init, one push and cc (whose scatter deposits the int64
    charge, read here), then charge_deposition and poisson_solver: the solve
    under test.  Returns the run, the rhs b of the interior rows (q 2^-32 /
    (ds eps0), exact in FP64) and the potential (FP64, with the Dirichlet
    rows), periodic layout."""
    p = cp.params(ncx=ncx, ncy=ncy, ppc=ppc, phi_p=-5.0, jacobi_mode=cp.POISSON_DIRECT,
                  jacobi_tol=1e-11, jacobi_tol_init=1e-11)
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


SHAPES = [
    (32, 16, 20), (4, 2, 4), (4, 8, 8), (1024, 4, 2), (256, 128, 2),   # FFT, lengths <= 1024
    (2048, 8, 1), (8, 1024, 1), (4096, 2, 1), (64, 1024, 1),         # FFT, lengths <= 4096
    (64, 30, 6), (48, 16, 6), (37, 9, 7), (100, 64, 2),               # dense products
]


@pytest.mark.parametrize("ncx,ncy,ppc", SHAPES, ids=[f"{a}x{b}" for a, b, _ in SHAPES])
def test_direct_solve_matches_sparse_solve(dev, ncx, ncy, ppc):
    run, b, phi = solved_run(ncx, ncy, ppc)
    sweeps, resid = run.last_solve()
    want = sparse_solve(b, phi, ncx, ncy)
    got = phi[1:ncy]
    err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
    say(f"{ncx}x{ncy}: FP64 residual {resid:.3e}, relative difference to the sparse solve {err:.3e}")
    assert resid < 1e-11
    assert err < 1e-11
    assert np.all(np.isfinite(phi))
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
    """A batch holding runs of both FFT instances and of the dense products
    steps each of them bit for bit as the run stepped alone."""
    shapes = [dict(ncx=64, ncy=32, ppc=6), dict(ncx=2048, ncy=8, ppc=2, seed=5),
              dict(ncx=40, ncy=20, ppc=6, seed=7), dict(ncx=16, ncy=1024, ppc=1, seed=9)]
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
