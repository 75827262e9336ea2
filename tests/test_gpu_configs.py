"""The benchmark configurations that build on the plain cycle, at their own
shapes, on the GPU through the C-ABI.

* Precision configuration (128x64 cells, 20 ppc, mi/me 100, phi_p -5):
  int64 fixed-point charge, FP32 push, FP64 probe-charge accumulation and
  Jacobi in FP32 with an FP64 residual and an FP64 correction every 50
  sweeps (refinement, at most 500 sweeps), run for 600 steps.  At
  checkpoints the error of rho, phi and E against FP64 on the GPU's own
  particles, of one step's particle states and probe counts against an
  FP64 step from the GPU's state, and over the run the probe current
  against an independent FP64 trajectory, are printed and bounded.
* Batched configuration: the bench's 256 runs (seed 20261203, grids
  64x32..512x256, 20..100 ppc, bias -10..+2, direct FP64 Poisson solve) in
  one HIP graph per step; every run's invariants, and 8 runs bit for bit against the same runs
  stepped alone.
* Serving configuration: the first requests of the bench's stream (grids
  128x64..1024x512, 20..200 ppc): each point's result is the same alone and
  in the mix, bit for bit.
"""
import ctypes as C
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


def errs(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = np.abs(want).max()
    d = np.abs(got - want)
    return d.max() / scale, float(np.sqrt((d * d).mean())) / scale


def fp64_fields(op, run, phi_start, sweeps):
    """rho, phi, Ex, Ey in FP64 from the run's particles: the FP64 deposit and
    the same number of FP64 Jacobi sweeps from the same starting potential
    (refinement equals plain Jacobi in exact arithmetic: the correction of k
    sweeps on the residual equation from zero adds up to k sweeps)."""
    nn = O.nodes(op)
    Q = np.zeros(nn)
    for s, q in ((0, -1.0), (1, 1.0)):
        x, y, _, _ = run.particles(s)
        O.lib().orc_deposit_f64(len(x), x.astype(np.float64), y.astype(np.float64), op.ds,
                                op.ncx, op.ncy, q, Q)
    rho = np.zeros(nn)
    O.lib().orc_rho_from_Q(op.ncx, op.ncy, op.ds, Q, rho)
    phi = np.ascontiguousarray(phi_start, np.float64).ravel().copy()
    work = np.zeros(nn)
    last = C.c_double()
    O.lib().orc_jacobi_sweeps(op.ncx, op.ncy, op.ds, op.eps0, rho, phi, work, sweeps,
                              C.byref(last))
    ex = np.zeros(nn)
    ey = np.zeros(nn)
    O.lib().orc_field(op.ncx, op.ncy, op.ds, phi, ex, ey)
    sh = (op.ncy + 1, op.ncx)
    return rho.reshape(sh), phi.reshape(sh), ex.reshape(sh), ey.reshape(sh)


def test_precision_configuration_run(dev):
    p = cp.params(ncx=128, ncy=64, ppc=20, mi=100.0, kti=1.0, phi_p=-5.0,
                  jacobi_mode=cp.JACOBI_REFINE, jacobi_check_every=50, jacobi_tol=1e-6,
                  jacobi_tol_init=1e-6, jacobi_max_sweeps=500)
    op = H.oracle_params(p)
    run = cp.Run(p)
    run.init_sim()
    sim = O.Sim(op, omp=True)        # an independent FP64 trajectory
    sim.set_solver(1, 1e-6, 1e-6, 0, 500, 50)
    sim.init()
    replay = O.Sim(op, omp=True)     # FP64 steps from the GPU's state
    replay.set_solver(1, 1e-6, 1e-6, 0, 500, 50)
    replay.init()
    steps, window0 = 600, 100
    checkpoints = (1, 200, 400, 600)
    fp64_probe = np.zeros(2, np.int64)
    gpu_window0 = None
    t = p.dt
    for step in range(1, steps + 1):
        if step in checkpoints:
            # the GPU state before this step, and an FP64 step from it
            pre = [run.particles(s) for s in (0, 1)]
            phi0 = H.periodic(run.mesh(cp.MESH_PHI))
            tin, ninj, started = run.clock()
            for s in (0, 1):
                replay.set_particles(s, *pre[s])
            replay.set_phi(phi0)
            replay.set_clock(tin, ninj, started)
            replay.step(t)
            # the GPU step, by entry points, with the solve's output compared
            run.charge_deposition()
            sweeps, resid = run.poisson_solver()
            rho64, phi64, ex64, ey64 = fp64_fields(op, run, phi0, sweeps)
            e_rho = errs(H.periodic(run.mesh(cp.MESH_RHO)), rho64)
            e_phi = errs(H.periodic(run.mesh(cp.MESH_PHI)), phi64)
            run.field_solver()
            e_ex = errs(H.periodic(run.mesh(cp.MESH_EX)), ex64)
            e_ey = errs(H.periodic(run.mesh(cp.MESH_EY)), ey64)
            run.particle_mover()
            run.cc(t)
            tally = run.probe_tally()
            rt = replay.tallies()
            # one step's particle states: FP32 against FP64 from the same start
            worst = 0.0
            from scipy.spatial import cKDTree
            for s in (0, 1):
                got = np.stack(run.particles(s), 1).astype(np.float64)
                want = np.stack(replay.particles(s), 1)
                assert len(got) == len(want), (step, s, len(got), len(want))
                scale = np.abs(want).max(0)
                dist, idx = cKDTree(want / scale).query(got / scale)
                assert len(np.unique(idx)) == len(got)
                worst = max(worst, float((np.abs(got - want[idx]).max(0) / scale).max()))
            say(f"step {step}: sweeps {sweeps} residual {resid:.2e} | rho max {e_rho[0]:.1e} "
                f"rms {e_rho[1]:.1e} | phi max {e_phi[0]:.1e} rms {e_phi[1]:.1e} | "
                f"Ex max {e_ex[0]:.1e} Ey max {e_ey[0]:.1e} | one-step states {worst:.1e} | "
                f"probe GPU {list(tally['probe'])} FP64 {list(rt['probe'])}")
            assert e_rho[0] <= 2e-5              # fixed-point weights, 2^-16 per axis
            assert e_phi[0] <= 1e-5 and e_ex[0] <= 1e-4 and e_ey[0] <= 1e-4
            assert worst <= 1e-5
            for s in (0, 1):
                assert abs(int(tally["probe"][s]) - int(rt["probe"][s])) <= 1
                assert tally["injected"][s] == rt["injected"][s]
        else:
            run.step(t)
        sim.step(t)
        fp64_probe += sim.tallies()["probe"]
        if step == window0:
            gpu_window0 = np.array(run.probe_tally()["total_probe"], np.int64)
            fp64_window0 = fp64_probe.copy()
        t += p.dt
    tally = run.probe_tally()
    gpu = np.array(tally["total_probe"], np.int64) - gpu_window0
    f64 = fp64_probe - fp64_window0
    # the FP64 probe charge is the exact sum of the absorbed charges
    assert tally["charge"] == float(int(tally["total_probe"][1]) - int(tally["total_probe"][0]))
    nsteps = steps - window0
    for s in (0, 1):
        a, b = int(gpu[s]), int(f64[s])
        rel = abs(a - b) / max(b, 1)
        noise = np.sqrt(a + b) / max(b, 1)
        say(f"species {s}: probe current over steps {window0 + 1}..{steps}: GPU "
            f"{a / (nsteps * p.dt):.2f}, FP64 {b / (nsteps * p.dt):.2f} particles per unit "
            f"time; difference {rel:.3%} (counting noise {noise:.3%})")
        # two independent trajectories: counts within 4 sigma of Poisson noise
        assert abs(a - b) <= 4.0 * np.sqrt(a + b) + 1
    run.close()


def draw_batched(n_runs, seed):
    """The batched benchmark's runs: ncx ~ U{64..512}, ncy ~ U{32..256},
    ppc ~ U{20..100}, bias ~ U[-10, 2] kTe/e."""
    rng = np.random.default_rng(seed)
    return [dict(ncx=int(rng.integers(64, 513)), ncy=int(rng.integers(32, 257)),
                 ppc=int(rng.integers(20, 101)), phi_p=float(rng.uniform(-10.0, 2.0)),
                 mi=100.0, kti=1.0, run_id=k) for k in range(n_runs)]


def batched_params(d):
    """The batched benchmark's solver: the direct FP64 solve to residual 1e-6."""
    return cp.params(**d, jacobi_mode=cp.POISSON_DIRECT, jacobi_tol=1e-6,
                     jacobi_tol_init=1e-6, jacobi_max_sweeps=500, jacobi_check_every=50)


def test_batched_configuration_256_runs(dev):
    draws = draw_batched(256, 20261203)
    runs = []
    for d in draws:
        r = cp.Run(batched_params(d))
        r.init_sim()
        runs.append(r)
    n0 = [[r.number_of_particles(s) for s in (0, 1)] for r in runs]
    say(f"256 runs, {sum(map(sum, n0))} particles")
    steps = 3
    b = cp.Batch(runs, graph=True)
    prev = n0
    for _ in range(steps):
        b.step(1)
        cur = []
        for k, r in enumerate(runs):
            tl = r.probe_tally()
            n1 = [r.number_of_particles(s) for s in (0, 1)]
            for s in (0, 1):
                # particle count: before - probe - top + injected, every run, every step
                assert n1[s] == prev[k][s] - tl["probe"][s] - tl["top"][s] + tl["injected"][s], k
            cur.append(n1)
        prev = cur
    b.close()
    for k, r in enumerate(runs):
        tally = r.probe_tally()
        n1 = prev[k]
        for s in (0, 1):
            off = r.offsets(s)
            assert off[0] == 0 and off[-1] == n1[s] and np.all(np.diff(off) >= 0), k
        q = r.mesh(cp.MESH_CHARGE)[:, :-1]
        assert int(q.sum()) == (n1[1] - n1[0]) * 2**32, k
        assert r.counters()["steps"] == steps
        assert tally["charge"] == float(int(tally["total_probe"][1]) -
                                        int(tally["total_probe"][0]))
    # 8 runs, the largest and smallest among them, stepped alone
    sizes = [sum(x) for x in n0]
    order = np.argsort(sizes)
    sample = [int(order[0]), int(order[-1])]
    for i in np.linspace(0, 255, 12).round().astype(int):
        if len(sample) < 8 and int(i) not in sample:
            sample.append(int(i))
    assert len(sample) == 8
    for k in sample:
        alone = cp.Run(batched_params(draws[k]))
        alone.init_sim()
        t = alone.p.dt
        for _ in range(steps):
            alone.step(t)
            t += alone.p.dt
        for s in (0, 1):
            a = np.stack(alone.particles(s))
            m = np.stack(runs[k].particles(s))
            assert a.tobytes() == m.tobytes(), (k, s)
            assert np.array_equal(alone.offsets(s), runs[k].offsets(s))
        assert np.array_equal(alone.mesh(cp.MESH_PHI), runs[k].mesh(cp.MESH_PHI))
        assert list(alone.probe_tally()["total_probe"]) == \
            list(runs[k].probe_tally()["total_probe"])
        say(f"run {k} ({draws[k]['ncx']}x{draws[k]['ncy']}, {draws[k]['ppc']} ppc): "
            f"alone == batched")
        alone.close()
    for r in runs:
        r.close()


def draw_requests(n, seed):
    """The serving benchmark's points: ncx ~ U{128..1024}, ncy ~ U{64..512},
    ppc ~ U{20..200}, bias ~ U[-10, 2] (arrival times drawn and unused)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        rng.exponential(1.0 / 50.0)
        out.append(dict(ncx=int(rng.integers(128, 1025)), ncy=int(rng.integers(64, 513)),
                        ppc=int(rng.integers(20, 201)), phi_p=float(rng.uniform(-10.0, 2.0)),
                        run_id=1000 + k))
    return out


SAME = ("status", "error", "steps", "windows", "current", "rel_change", "total_probe",
        "particles")


def test_serving_configuration_alone_and_in_mix(dev):
    reqs = draw_requests(4, 20261203)
    cfg = dict(round_steps=20, window_steps=100, max_steps=200, rel_change=0.005)

    def drain(srv, ids):
        for _ in range(1000):
            active, queued = srv.poll()
            if active == 0 and queued == 0:
                break
        return [srv.result(i) for i in ids]

    alone = []
    for d in reqs:
        srv = cp.Server(cp.server_config(**cfg))
        alone.append(drain(srv, [srv.submit(cp.params(**d))])[0])
        srv.close()
        say(f"alone {d['ncx']}x{d['ncy']} {d['ppc']} ppc: "
            f"{ {k: alone[-1][k] for k in ('status', 'steps', 'total_probe')} }")
    srv = cp.Server(cp.server_config(**cfg))
    ids = [srv.submit(cp.params(**reqs[0])), srv.submit(cp.params(**reqs[1]))]
    srv.poll()                               # the others join a running batch
    ids += [srv.submit(cp.params(**d)) for d in reqs[2:]]
    mixed = drain(srv, ids)
    st = srv.stats()
    srv.close()
    say("mix: rounds", st["rounds"], "rebuilds", st["rebuilds"])
    for a, m in zip(alone, mixed):
        assert a["status"] in (cp.POINT_CONVERGED, cp.POINT_CAPPED)
        for k in SAME:
            assert m[k] == a[k], (k, m[k], a[k])
