"""Error handling on the GPU path, through the C-ABI: every failure is a
status code with a message (the reference prints and exits, src/cuda.cu:15-54),
the kernels' own bounds checks turn overflows into device flags, and no
kernel runs out of bounds.  Conditions are forced from the host."""
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


def test_far_list_overflow_is_reported(dev):
    """dt = 4 on a 64 x 128 grid at 40 particles per cell: about 60 % of the
    electrons cross more than one cell in one step, roughly three times what
    the far list holds (its capacity follows the injection rate); the step
    completes and the next call that reads the state reports
    CUPIC_ERR_CAPACITY."""
    p = cp.params(ncx=64, ncy=128, ppc=40, dt=4.0, jacobi_max_sweeps=50)
    run = cp.Run(p)
    run.init_sim()
    say("far-list overflow step")
    run.step(p.dt)
    with pytest.raises(cp.CupicError) as e:
        run.probe_tally()
    assert e.value.code == -4 and "far-list overflow" in str(e.value)
    run.close()


def test_calls_out_of_order(dev):
    p = cp.params(ncx=32, ncy=16, ppc=4)
    run = cp.Run(p)
    run.init_sim()
    with pytest.raises(cp.CupicError) as e:       # nothing pending: charge was consumed
        run.charge_deposition()
        run.mesh(cp.MESH_CHARGE)
    assert e.value.code == -5
    run.particle_mover()
    with pytest.raises(cp.CupicError) as e:       # deposit between mover and cc
        run.charge_deposition()
    assert e.value.code == -5
    with pytest.raises(cp.CupicError) as e:
        run.set_particles(0, *[np.zeros(3, np.float32)] * 4)
    assert e.value.code == -5
    run.cc(p.dt)                                  # the pending cc still completes
    assert run.number_of_particles(0) > 0
    run.close()


def test_capacity_limits(dev):
    p = cp.params(ncx=32, ncy=16, ppc=4)
    n = p.per_row * p.ncy
    with pytest.raises(cp.CupicError) as e:       # below the initial load
        cp.Run(p, capacity=n - 1)
    assert e.value.code == -1
    run = cp.Run(p, capacity=n + 16)
    run.init_sim()
    x = np.full(n + 17, 1.0, np.float32)
    with pytest.raises(cp.CupicError) as e:
        run.set_particles(0, x, x, x, x)
    assert e.value.code == -4
    # within capacity it works and the state is sorted again
    run.set_particles(0, x[:10], x[:10], x[:10], x[:10])
    assert run.number_of_particles(0) == 10
    run.close()


def test_bad_arguments(dev):
    run = cp.Run(cp.params(ncx=32, ncy=16, ppc=4))
    buf = np.zeros(33 * 17, np.float64)
    assert cp.lib.cupic_get_mesh(run.h, 9, buf.ctypes.data, buf.size) == -1
    assert cp.lib.cupic_get_mesh(run.h, cp.MESH_PHI, buf.ctypes.data, buf.size - 1) == -1
    for bad in (lambda: run.particles_snapshot(2, "/tmp/x", 0.0), lambda: run.set_tile(17)):
        with pytest.raises(cp.CupicError) as e:
            bad()
        assert e.value.code == -1
    with pytest.raises(cp.CupicError) as e:
        run.particles_snapshot(0, "/nonexistent-dir/electrons", 0.0)
    assert e.value.code == -7
    run.close()
