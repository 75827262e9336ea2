"""The no-allocation Poisson solve (cupic_poisson_solver_noalloc) against the
run's own solve (cupic_poisson_solver) on the same node charge and starting
potential: bit-identical potential, rho, sweep count and residual, on calls
in a row over shrinking grids (the scratch keeps the previous call's
contents), in both stopping rules.  The caller's buffers are plain device
allocations made here with the HIP runtime; the entry point allocates
nothing itself."""
import ctypes as C
import sys

import numpy as np
import pytest

import helpers as H

cp = pytest.importorskip("cupic_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DevBuf:
    """A device buffer from the HIP runtime the library itself uses."""
    hip = None

    def __init__(self, nbytes):
        if DevBuf.hip is None:
            DevBuf.hip = C.CDLL("libamdhip64.so.7")
            DevBuf.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            DevBuf.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            DevBuf.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        assert DevBuf.hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert DevBuf.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def download(self, dtype, shape):
        a = np.empty(shape, dtype)
        assert a.nbytes == self.nbytes
        assert DevBuf.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, 2) == 0
        return a

    def free(self):
        if self.ptr:
            DevBuf.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def test_noalloc_rejects_what_the_scratch_cannot_hold():
    """Checked on the host before any device call (runs without a GPU)."""
    big = cp.params(ncx=2048, ncy=1024, ppc=1)
    with pytest.raises(cp.CupicError) as e:
        cp.poisson_solver_noalloc(big, 1, 1)
    assert e.value.code == -1 and "CUPIC_NOALLOC_SCRATCH_BYTES" in str(e.value)
    many = cp.params(jacobi_max_sweeps=200000)
    with pytest.raises(cp.CupicError) as e:
        cp.poisson_solver_noalloc(many, 1, 1)
    assert e.value.code == -1
    with pytest.raises(cp.CupicError) as e:
        cp.poisson_solver_noalloc(cp.params(jacobi_mode=cp.POISSON_DIRECT), 1, 1)
    assert e.value.code == -1 and "jacobi_mode" in str(e.value)


CASES = [
    # production grid, FP64-residual rule, 500-sweep cap (one launch runs
    # through the residual checks)
    dict(ncx=1024, ncy=512, ppc=2, mi=1836.0, kti=0.1, phi_p=-10.0,
         jacobi_mode=cp.JACOBI_RESIDUAL, jacobi_tol=1e-6, jacobi_max_sweeps=500),
    # smaller, loose tolerance: the solve stops at a check inside a launch
    dict(ncx=128, ncy=64, ppc=20, jacobi_mode=cp.JACOBI_RESIDUAL, jacobi_tol=2e-4,
         jacobi_max_sweeps=500),
    # smallest, the reference rule (max |dphi| after a minimum of sweeps)
    dict(ncx=37, ncy=9, ppc=7, phi_p=2.0, jacobi_mode=cp.JACOBI_REFERENCE, jacobi_tol=1e-4,
         jacobi_min_sweeps=38, jacobi_max_sweeps=400),
]


@pytest.mark.gpu
def test_noalloc_solver_equals_run_solver():
    name, _ = cp.init_dev(0)
    say("device:", name)
    for kw in CASES:
        p = cp.params(**kw)
        run = cp.Run(p)
        run.init_sim()
        run.step(p.dt)          # leaves the fused charge of the moved particles pending
        Q = H.periodic(run.mesh(cp.MESH_CHARGE)).astype(np.int64)
        phi0 = H.periodic(run.mesh(cp.MESH_PHI)).astype(np.float32)
        shape = Q.shape
        dQ = DevBuf(Q.nbytes)
        dphi = DevBuf(phi0.nbytes)
        drho = DevBuf(phi0.nbytes)
        dQ.upload(Q)
        dphi.upload(phi0)
        drho.upload(np.full(shape, np.nan, np.float32))
        # the run's own solve on the same charge and starting potential
        run.charge_deposition()
        sw_ref, res_ref = run.poisson_solver()
        phi_ref = H.periodic(run.mesh(cp.MESH_PHI)).astype(np.float32)
        rho_ref = H.periodic(run.mesh(cp.MESH_RHO)).astype(np.float32)
        say(f"{kw['ncx']}x{kw['ncy']} mode {kw['jacobi_mode']}: run solve {sw_ref} sweeps, "
            f"residual {res_ref:.3g}")
        sw, res = cp.poisson_solver_noalloc(run.p, dQ.ptr.value, dphi.ptr.value,
                                            drho.ptr.value)
        say(f"   no-allocation solve {sw} sweeps, residual {res:.3g}")
        phi = dphi.download(np.float32, shape)
        rho = drho.download(np.float32, shape)
        assert (sw, res) == (sw_ref, res_ref)
        assert phi.tobytes() == phi_ref.tobytes(), H.mismatch(phi, phi_ref)
        assert rho.tobytes() == rho_ref.tobytes(), H.mismatch(rho, rho_ref)
        assert np.array_equal(dQ.download(np.int64, shape), Q)   # the charge is read only
        # the Dirichlet rows are the caller's
        assert np.array_equal(phi[0], phi0[0]) and np.array_equal(phi[-1], phi0[-1])
        for b in (dQ, dphi, drho):
            b.free()
        run.close()
