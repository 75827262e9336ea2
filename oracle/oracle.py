"""ctypes bindings for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The shared libraries are built by
oracle/Makefile (``make -C oracle``).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

Q7_REFERENCE_KICK = 1


class Params(C.Structure):
    _fields_ = [
        ("ncx", C.c_int32), ("ncy", C.c_int32),
        ("ds", C.c_double), ("dt", C.c_double), ("n", C.c_double),
        ("mi", C.c_double), ("kti", C.c_double), ("me", C.c_double), ("kte", C.c_double),
        ("phi_p", C.c_double), ("eps0", C.c_double),
        ("seed", C.c_uint64), ("run_id", C.c_uint64),
        ("quirks", C.c_uint32), ("pad0", C.c_int32),
        ("Lx", C.c_double), ("Ly", C.c_double),
        ("dtin_e", C.c_double), ("dtin_i", C.c_double),
        ("per_row", C.c_int64),
    ]


class F32C(C.Structure):
    _fields_ = [
        ("inv_ds", C.c_float), ("Lx", C.c_float), ("Ly", C.c_float), ("dt", C.c_float),
        ("qdtm", C.c_float * 2), ("hkick", C.c_float * 2),
        ("ncx", C.c_int32), ("ncy", C.c_int32),
    ]


def _load(omp=False):
    name = "libcupic_oracle_omp.so" if omp else "libcupic_oracle.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise OSError(f"oracle library {path} is not built (run: make -C oracle)")
    lib = C.CDLL(path)
    P = C.POINTER(Params)
    F = C.POINTER(F32C)
    vp = C.c_void_p
    sig = {
        "orc_params_derive": (None, [P]),
        "orc_philox4x64_10": (None, [u64p, u64p, u64p]),
        "orc_u01_open": (C.c_double, [C.c_uint64]),
        "orc_draw4": (None, [P, C.c_int, C.c_int, C.c_uint64, u64p]),
        "orc_create_particles_f64": (None, [P, C.c_int, f64p, f64p, f64p, f64p]),
        "orc_deposit_f64": (None, [C.c_int64, f64p, f64p, C.c_double, C.c_int, C.c_int,
                                   C.c_double, f64p]),
        "orc_rho_from_Q": (None, [C.c_int, C.c_int, C.c_double, f64p, f64p]),
        "orc_phi_linear": (None, [C.c_int, C.c_int, C.c_double, f64p]),
        "orc_jacobi_sweeps": (None, [C.c_int, C.c_int, C.c_double, C.c_double, f64p, f64p,
                                     f64p, C.c_int, C.POINTER(C.c_double)]),
        "orc_residual_max": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, f64p,
                                          f64p]),
        "orc_poisson_solve": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, f64p, f64p,
                                        C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_double)]),
        "orc_field": (None, [C.c_int, C.c_int, C.c_double, f64p, f64p, f64p]),
        "orc_push_f64": (None, [C.c_int64, f64p, f64p, f64p, f64p, f64p, f64p, C.c_int,
                                C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
        "orc_half_kick_f64": (None, [C.c_int64, f64p, f64p, f64p, f64p, f64p, f64p, C.c_int,
                                     C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double]),
        "orc_inject_f64": (None, [P, C.c_int, C.c_uint64, C.c_int, C.c_double, f64p, f64p,
                                  f64p, f64p, f64p, f64p, f64p]),
        "orc_injection_count": (C.c_int, [C.c_double, C.c_double, C.POINTER(C.c_double),
                                          C.c_double, vp, C.c_int]),
        "orc_sim_create": (vp, [P, C.c_int64]),
        "orc_sim_destroy": (None, [vp]),
        "orc_sim_set_solver": (None, [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int]),
        "orc_sim_init": (None, [vp]),
        "orc_sim_step": (None, [vp, C.c_double]),
        "orc_sim_count": (C.c_int64, [vp, C.c_int]),
        "orc_sim_get_particles": (None, [vp, C.c_int, f64p, f64p, f64p, f64p]),
        "orc_sim_set_particles": (None, [vp, C.c_int, C.c_int64, f64p, f64p, f64p, f64p]),
        "orc_sim_get_mesh": (None, [vp, f64p, f64p, f64p, f64p]),
        "orc_sim_set_phi": (None, [vp, f64p]),
        "orc_sim_tallies": (None, [vp, i64p, i64p, i64p, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double)]),
        "orc_sim_clock": (None, [vp, f64p, u64p]),
        "orc_sim_set_clock": (None, [vp, f64p, u64p, C.c_int]),
        "orc_f32c_make": (None, [P, F]),
        "orc_cell_of_f32": (C.c_int32, [F, C.c_float, C.c_float, C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]),
        "orc_create_particles_f32": (None, [P, C.c_int, f32p, f32p, f32p, f32p]),
        "orc_deposit_fixed_f32": (None, [C.c_int64, f32p, f32p, F, C.c_int, i64p]),
        "orc_stable_sort_slots": (C.c_int64, [C.c_int64, i32p, i8p, C.c_int32, i64p, i64p]),
        "orc_mirror_init": (None, [P, C.c_int, f32p, f32p, f32p, f32p, i64p]),
        "orc_mirror_half_kick": (None, [F, C.c_int, C.c_int64, f32p, f32p, f32p, f32p, f32p]),
        "orc_mirror_step": (C.c_int64, [F, C.c_int, C.c_int64, f32p, f32p, f32p, f32p, f32p,
                                        C.c_int64, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                        f32p, i64p, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "orc_inject_f32": (None, [P, C.c_int, C.c_uint64, C.c_int, C.c_double, f64p, f32p,
                                  f32p, f32p, f32p, f32p]),
        "orc_mirror_rhs": (None, [P, i64p, f64p, f32p]),
        "orc_mirror_solve_f32": (C.c_int, [C.c_int, C.c_int, f64p, f32p, f32p, C.c_double,
                                           C.c_int, C.c_int, C.POINTER(C.c_double)]),
    }
    for k, (res, args) in sig.items():
        fn = getattr(lib, k)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIBS = {}


def lib(omp=False):
    if omp not in _LIBS:
        _LIBS[omp] = _load(omp)
    return _LIBS[omp]


def make_params(ncx=128, ncy=64, ds=0.5, dt=0.05, ppc=20, mi=100.0, kti=1.0, phi_p=-5.0,
                seed=20261203, run_id=0, quirks=0):
    """Config-1 defaults; n chosen so that n*ds^3 = ppc particles per cell."""
    p = Params()
    p.ncx, p.ncy, p.ds, p.dt = ncx, ncy, ds, dt
    p.n = ppc / (ds * ds * ds)
    p.mi, p.kti, p.me, p.kte = mi, kti, 1.0, 1.0
    p.phi_p, p.eps0 = phi_p, 0.0
    p.seed, p.run_id, p.quirks = seed, run_id, quirks
    lib().orc_params_derive(C.byref(p))
    return p


def f32c(p):
    c = F32C()
    lib().orc_f32c_make(C.byref(p), C.byref(c))
    return c


def nodes(p):
    return p.ncx * (p.ncy + 1)


class Sim:
    """FP64 long-run oracle (src/main.cu:59-96 cycle)."""

    def __init__(self, p, capacity=0, omp=False):
        self.L = lib(omp)
        self.p = p
        self.h = self.L.orc_sim_create(C.byref(p), capacity)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_sim_destroy(self.h)
            self.h = None

    def set_solver(self, mode, tol, tol_init, min_sweeps, max_sweeps, check_every=50):
        self.L.orc_sim_set_solver(self.h, mode, tol, tol_init, min_sweeps, max_sweeps,
                                  check_every)

    def init(self):
        self.L.orc_sim_init(self.h)

    def step(self, t):
        self.L.orc_sim_step(self.h, t)

    def count(self, s):
        return self.L.orc_sim_count(self.h, s)

    def particles(self, s):
        n = self.count(s)
        a = [np.empty(n, np.float64) for _ in range(4)]
        self.L.orc_sim_get_particles(self.h, s, *a)
        return a

    def set_particles(self, s, x, y, vx, vy):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, vx, vy)]
        self.L.orc_sim_set_particles(self.h, s, len(arrs[0]), *arrs)

    def mesh(self):
        m = [np.empty(nodes(self.p), np.float64) for _ in range(4)]
        self.L.orc_sim_get_mesh(self.h, *m)
        sh = (self.p.ncy + 1, self.p.ncx)
        return [v.reshape(sh) for v in m]

    def set_phi(self, phi):
        self.L.orc_sim_set_phi(self.h, np.ascontiguousarray(phi, dtype=np.float64).ravel())

    def tallies(self):
        pr = np.zeros(2, np.int64)
        tp = np.zeros(2, np.int64)
        inj = np.zeros(2, np.int64)
        sw = C.c_int(0)
        res = C.c_double(0)
        self.L.orc_sim_tallies(self.h, pr, tp, inj, C.byref(sw), C.byref(res))
        return dict(probe=pr, top=tp, injected=inj, sweeps=sw.value, resid=res.value)

    def clock(self):
        tin = np.zeros(2, np.float64)
        ninj = np.zeros(2, np.uint64)
        self.L.orc_sim_clock(self.h, tin, ninj)
        return tin, ninj

    def set_clock(self, tin, ninj, started=1):
        self.L.orc_sim_set_clock(self.h, np.asarray(tin, np.float64),
                                 np.asarray(ninj, np.uint64), started)
