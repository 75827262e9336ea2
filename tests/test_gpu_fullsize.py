"""Size-independent properties at the production size (1024x512 cells, 200
particles per cell per species, 2 x 104,857,600 particles), where the oracle
is too slow to replay every step: per step, N' = N - probe - top + injected
for each species; the fixed-point node charge sums to (Ni - Ne) 2^32
exactly; every particle is inside the domain; the state is sorted by cell
key and the offsets are the exclusive scan of the cell histogram.  The
headline configuration is also run for its whole 5000 steps."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

cp = pytest.importorskip("cupic_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_production_size_invariants():
    cp.init_dev(0)
    p = cp.params(ncx=1024, ncy=512, ppc=200, mi=1836.0, kti=0.1, phi_p=-10.0,
                  jacobi_mode=cp.JACOBI_RESIDUAL, jacobi_tol=1e-6, jacobi_tol_init=1e-6,
                  jacobi_max_sweeps=500, jacobi_check_every=50)
    say("production size: 1024x512, 200 ppc")
    run = cp.Run(p)
    run.init_sim()
    n = [run.number_of_particles(s) for s in (0, 1)]
    assert n == [p.per_row * p.ncy] * 2
    t = p.dt
    for step in range(3):
        run.step(t)
        t += p.dt
        tally = run.probe_tally()
        new = [run.number_of_particles(s) for s in (0, 1)]
        for s in (0, 1):
            assert new[s] == n[s] - tally["probe"][s] - tally["top"][s] + tally["injected"][s]
        charge = run.mesh(cp.MESH_CHARGE)[:, :-1]
        assert int(charge.sum()) == (new[1] - new[0]) * (1 << 32)
        sweeps, resid = run.last_solve()
        say(f"  step {step}: N {new}, probe {list(tally['probe'])}, "
            f"injected {list(tally['injected'])}, sweeps {sweeps}, residual {resid:.3g}")
        n = new
    Lx, Ly, ds = p.ncx * p.ds, p.ncy * p.ds, p.ds
    for s in (0, 1):
        x, y, vx, vy = run.particles(s)
        assert len(x) == n[s]
        assert np.all((x >= 0) & (x < Lx) & (y >= 0) & (y < Ly))
        assert np.all(np.isfinite(vx)) and np.all(np.isfinite(vy))
        ic = np.minimum((x / np.float32(ds)).astype(np.int64), p.ncx - 1)
        jc = np.minimum((y / np.float32(ds)).astype(np.int64), p.ncy - 1)
        key = jc * p.ncx + ic
        del x, y, vx, vy
        assert np.all(np.diff(key) >= 0)
        off = run.offsets(s).astype(np.int64)
        hist = np.bincount(key, minlength=p.ncx * p.ncy)
        assert off[0] == 0 and np.array_equal(np.diff(off), hist)
        say(f"  species {s}: sorted, offsets match the histogram")
    run.close()


def test_production_configuration_whole_run():
    """The headline configuration over its whole 5000-step run, as the bench
    times it (the Poisson system solved directly in FP64): particle counts
    conserved every step (N' = N - probe - top + injected), the node charge
    exact and the FP64 residual below the configuration's 1e-6 every 250
    steps, the potential between the probe bias and 20 kTe/e throughout
    (it oscillates, up to 12 kTe/e), and the final state inside the domain
    and sorted by cell."""
    cp.init_dev(0)
    p = cp.params(ncx=1024, ncy=512, ppc=200, mi=1836.0, kti=0.1, phi_p=-10.0,
                  jacobi_mode=cp.POISSON_DIRECT, jacobi_tol=1e-6, jacobi_tol_init=1e-6)
    say("production configuration, 5000 steps, direct FP64 solve")
    run = cp.Run(p)
    run.init_sim()
    n = [run.number_of_particles(s) for s in (0, 1)]
    t = p.dt
    lo, hi = 0.0, 0.0
    for step in range(1, 5001):
        run.step(t)
        t += p.dt
        tally = run.probe_tally()
        new = [run.number_of_particles(s) for s in (0, 1)]
        for s in (0, 1):
            assert new[s] == n[s] - tally["probe"][s] - tally["top"][s] + tally["injected"][s], \
                (step, s)
        n = new
        if step % 250 == 0:
            sweeps, resid = run.last_solve()
            phi = run.mesh(cp.MESH_PHI)
            charge = run.mesh(cp.MESH_CHARGE)[:, :-1]
            assert int(charge.sum()) == (n[1] - n[0]) * (1 << 32), step
            lo, hi = min(lo, float(phi.min())), max(hi, float(phi.max()))
            say(f"  step {step}: N {n}, residual {resid:.2e}, phi {phi.min():.3f}..{phi.max():.3f}, "
                f"probe totals {list(tally['total_probe'])}")
            assert resid < 1e-6, (step, resid)
            # the plasma potential oscillates (peaks of 5-12 kTe/e about every
            # 750 steps, troughs below 1); a runaway such as the capped Jacobi
            # solve's (-198..+152 kTe/e by step 500) is far outside this band
            assert phi.min() >= p.phi_p - 1.0 and phi.max() <= 20.0, (step, phi.min(), phi.max())
    say(f"  potential over the run: {lo:.3f}..{hi:.3f} kTe/e")
    Lx, Ly, ds = p.ncx * p.ds, p.ncy * p.ds, p.ds
    for s in (0, 1):
        x, y, vx, vy = run.particles(s)
        assert len(x) == n[s]
        assert np.all((x >= 0) & (x < Lx) & (y >= 0) & (y < Ly))
        assert np.all(np.isfinite(vx)) and np.all(np.isfinite(vy))
        ic = np.minimum((x / np.float32(ds)).astype(np.int64), p.ncx - 1)
        jc = np.minimum((y / np.float32(ds)).astype(np.int64), p.ncy - 1)
        del x, y, vx, vy
        assert np.all(np.diff(jc * p.ncx + ic) >= 0)
        say(f"  species {s}: {n[s]} particles, inside the domain and sorted")
    run.close()
