"""Long runs at the small configuration (128x64 cells, 20 particles per cell
per species, mi/me 100, Te/Ti 1, probe at -5 kTe/e, dt 0.05, Jacobi with
FP64 residual 1e-6 and at most 500 sweeps):

* against the oracle's FP32 mirror of the whole cycle (its own deposit,
  FP32 Jacobi, field, push, injection and sort; no GPU data enters it):
  bit-identical particle states, potential, probe current and density
  profiles after hundreds of steps;
* against the FP64 oracle (an independent trajectory, since FP32 and FP64
  round differently and the dynamics is chaotic): the time-averaged density
  profiles agree within 1 % and the collected probe charge within its
  counting noise.
"""
import sys

import numpy as np
import pytest

import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

cp = pytest.importorskip("cupic_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    return cp.init_dev(0)


def test_long_run_bit_exact_against_fp32_mirror(dev):
    p = cp.params()
    run = cp.Run(p)
    run.init_sim()
    m = H.MirrorRun(p)
    t = p.dt
    for step in range(1, 401):
        run.step(t)
        m.step(t)
        t += p.dt
        if step in (1, 100, 200, 400):
            sweeps, resid = run.last_solve()
            tally = run.probe_tally()
            say(f"step {step}: N {[len(s[0]) for s in m.st]} probe {m.total_probe} "
                f"sweeps {sweeps}/{m.sweeps} residual {resid:.3g}/{m.resid:.3g}")
            assert sweeps == m.sweeps and resid == m.resid
            assert list(tally["total_probe"]) == m.total_probe
            phi = H.periodic(run.mesh(cp.MESH_PHI)).astype(np.float32).ravel()
            assert phi.tobytes() == m.phi.tobytes()
            for s in (0, 1):
                got = np.stack(run.particles(s))
                assert got.tobytes() == np.stack(m.st[s]).tobytes(), (step, s)
                assert np.array_equal(run.offsets(s), m.off[s].astype(np.int32))
                assert np.array_equal(run.density_profile(s), m.row_profile(s))
    run.close()


def coarse(profile, bins):
    return profile.reshape(bins, -1).sum(axis=1)


def test_long_run_statistics_against_fp64_oracle(dev):
    p = cp.params()
    steps, first, every, bins = 1000, 500, 10, 8
    run = cp.Run(p)
    run.init_sim()
    op = H.oracle_params(p)
    sim = O.Sim(op, omp=True)
    sim.set_solver(1, p.jacobi_tol, p.jacobi_tol_init, 0, p.jacobi_max_sweeps,
                   p.jacobi_check_every)
    sim.init()
    gpu_prof = np.zeros((2, bins))
    fp64_prof = np.zeros((2, bins))
    fp64_probe = np.zeros(2, np.int64)
    t = p.dt
    samples = 0
    for step in range(1, steps + 1):
        run.step(t)
        sim.step(t)
        t += p.dt
        fp64_probe += sim.tallies()["probe"]
        if step >= first and step % every == 0:
            samples += 1
            for s in (0, 1):
                gpu_prof[s] += coarse(run.density_profile(s), bins)
                y = sim.particles(s)[1]
                rows = np.minimum((y / op.ds).astype(np.int64), op.ncy - 1)
                fp64_prof[s] += coarse(np.bincount(rows, minlength=op.ncy), bins)
    gpu_probe = run.probe_tally()["total_probe"]
    rel = np.abs(gpu_prof - fp64_prof) / fp64_prof
    say("density profile (8 bins, steps 500..1000), GPU/FP64 - 1:\n", np.round(rel, 5))
    say("probe counts GPU", list(gpu_probe), "FP64", list(fp64_probe))
    assert rel.max() < 0.01
    for s in (0, 1):
        # two independent Poisson counts: |a - b| within 4 sigma of sqrt(a + b)
        a, b = int(gpu_probe[s]), int(fp64_probe[s])
        assert abs(a - b) <= 4.0 * np.sqrt(a + b) + 1, (s, a, b)
    run.close()
