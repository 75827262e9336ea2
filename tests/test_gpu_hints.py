"""Push hints of the counting pass change speed only.  Runs stepped with the
hints off, on and checked hold bit-identical particles, charge and probe
tallies after every block of steps, and the checked mode (every trusted hint
compared with the push that replaces it) raises no device error: on the
small configuration, under a probe bias strong enough that the field
overturns many hints, with a time step that moves electrons across cells
and into the far list often, and at full size (1024 x 512 cells, 200
particles per cell and species)."""
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


def same_state(a, b):
    for s in (0, 1):
        for u, v in zip(a.particles(s), b.particles(s)):
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    assert np.array_equal(a.mesh(cp.MESH_RHO).view(np.uint32), b.mesh(cp.MESH_RHO).view(np.uint32))
    ta, tb = a.probe_tally(), b.probe_tally()
    for k in ("probe", "top", "injected", "total_probe"):
        assert np.array_equal(ta[k], tb[k]), k


CASES = {
    "small": dict(ncx=128, ncy=64, ppc=20, phi_p=-5.0),
    "strong-bias": dict(ncx=96, ncy=48, ppc=30, phi_p=-400.0),
    "long-dt": dict(ncx=64, ncy=64, ppc=20, dt=0.2, phi_p=-20.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_hint_modes_identical(dev, name):
    kw = dict(CASES[name], jacobi_mode=cp.POISSON_DIRECT)
    say(f"hint modes: {name} {kw}")
    runs = []
    for mode in (cp.HINTS_OFF, cp.HINTS_ON, cp.HINTS_CHECKED):
        p = cp.params(**kw)
        r = cp.Run(p)
        r.set_hint_mode(mode)
        r.init_sim()
        runs.append(r)
    t = p.dt
    for i in range(1, 121):
        for r in runs:
            r.step(t)
        t += p.dt
        if i % 40 == 0:
            runs[2].probe_tally()  # raises on a contradicted hint
            same_state(runs[0], runs[1])
            same_state(runs[0], runs[2])
    for r in runs:
        r.close()


def test_hints_checked_full_size(dev):
    p = cp.params(ncx=1024, ncy=512, ppc=200, mi=1836.0, kti=0.1, phi_p=-10.0,
                  jacobi_mode=cp.POISSON_DIRECT)
    say("hints checked: 1024 x 512, 200 ppc, 10 steps")
    r = cp.Run(p)
    r.set_hint_mode(cp.HINTS_CHECKED)
    r.init_sim()
    t = p.dt
    for i in range(10):
        r.step(t)
        t += p.dt
    tally = r.probe_tally()  # raises on a contradicted hint
    say(f"probe {tally['total_probe']}")
    r.close()


def test_hint_mode_rejects_bad_values(dev):
    r = cp.Run(cp.params(ncx=16, ncy=8, ppc=4))
    with pytest.raises(cp.CupicError):
        r.set_hint_mode(3)
    r.close()
