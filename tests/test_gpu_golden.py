"""GPU path against the committed golden fixtures (tests/golden/*.npz), through
the C-ABI: initial load and sort, offsets, the field of a given potential,
the half-step stagger, one push/absorb/inject/sort step and its fixed-point
charge (all bit-exact), and the potential that initialize_mesh solves from
the initial charge (FP64 iterative refinement within 1e-8 relative, the
direct FP64 solve within 1e-11)."""
import glob
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

cp = pytest.importorskip("cupic_amd")

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "golden_*.npz")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def params_of(d, **kw):
    base = dict(ncx=int(d["ncx"]), ncy=int(d["ncy"]), ds=float(d["ds"]), dt=float(d["dt"]),
                n=float(d["n"]), mi=float(d["mi"]), kti=float(d["kti"]),
                phi_p=float(d["phi_p"]), seed=int(d["seed"]), run_id=int(d["run_id"]))
    base.update(kw)
    p = cp.params(**base)
    assert p.per_row == int(d["per_row"])
    return p


def reference_layout(a):
    """periodic (nny, ncx) -> reference (nny, nnx) with the image column"""
    return np.concatenate([a, a[:, :1]], axis=1)


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.fixture(scope="module")
def dev():
    return cp.init_dev(0)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_golden_step_chain(dev, path):
    d = load(path)
    say("golden chain", os.path.basename(path))
    run = cp.Run(params_of(d))
    run.create_particles()
    for s in (0, 1):
        got = np.stack(run.particles(s))
        assert same_bits(got, d[f"init_{s}"]), ("init", s)
        assert np.array_equal(run.offsets(s), d[f"off_{s}"]), ("offsets", s)
    run.set_phi(reference_layout(d["phi32"].astype(np.float64)))
    run.field_solver()
    ex = run.mesh(cp.MESH_EX)[:, :-1]
    ey = run.mesh(cp.MESH_EY)[:, :-1]
    assert same_bits(ex, d["E"][0::2].reshape(ex.shape))
    assert same_bits(ey, d["E"][1::2].reshape(ey.shape))
    run.adjust_leap_frog()
    for s in (0, 1):
        got = run.particles(s)
        assert same_bits(np.stack(got[2:]), d[f"kick_{s}"]), ("kick", s)
    run.particle_mover()
    run.cc(float(d["dt"]))
    tally = run.probe_tally()
    tin, _, started = run.clock()
    assert started
    for s in (0, 1):
        got = np.stack(run.particles(s))
        assert same_bits(got, d[f"step_{s}"]), ("step", s)
        assert np.array_equal(run.offsets(s), d[f"off1_{s}"]), ("offsets after step", s)
        pr, tp, inj = d[f"counts_{s}"]
        assert (tally["probe"][s], tally["top"][s], tally["injected"][s]) == (pr, tp, inj)
        assert tin[s] == float(d[f"tin_{s}"])
    charge = run.mesh(cp.MESH_CHARGE)[:, :-1]
    assert np.array_equal(charge, d["charge1"])
    run.close()


@pytest.mark.parametrize("mode,tol", [(cp.JACOBI_REFINE, 1e-8), (cp.POISSON_DIRECT, 1e-11)],
                         ids=["refine", "direct"])
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_golden_charge_and_potential(dev, path, mode, tol):
    """Iterative refinement to an FP64 residual of 1e-11, or the direct FP64
    solve of the same system, against the fixture's converged potential."""
    d = load(path)
    say("golden potential", os.path.basename(path), "mode", mode)
    p = params_of(d, jacobi_mode=mode, jacobi_max_sweeps=20000,
                  jacobi_check_every=50, jacobi_tol=1e-11, jacobi_tol_init=1e-11)
    run = cp.Run(p)
    run.create_particles()
    # linear potential with the Dirichlet rows, deposit, solve, field
    # (src/init.cu:202-260), here to the FP64 tolerance 1e-11
    run.initialize_mesh()
    # rho = Q 2^-32 / V with V = ds^3 (ds^3/2 on the Dirichlet rows); exact in binary
    ds3 = float(d["ds"]) ** 3
    q = d["charge0"].astype(np.float64) * 2.0 ** -32
    vol = np.full(q.shape[0], ds3)
    vol[0] = vol[-1] = 0.5 * ds3
    rho = run.mesh(cp.MESH_RHO)[:, :-1]
    assert np.array_equal(rho, (q / vol[:, None]).astype(np.float32))
    sweeps, resid = run.last_solve()
    say("  sweeps", sweeps, "residual", resid)
    assert resid < 1e-11
    phi = run.mesh(cp.MESH_PHI)[:, :-1]
    want = d["phi"]
    err = np.abs(phi - want).max() / np.abs(want).max()
    say("  potential relative error", err)
    assert err < tol
    run.close()
