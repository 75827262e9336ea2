"""cupic_amd -- MI355X-native electrostatic PIC cycle (cupic's hot path).

Python mirror of the reference's module entry points over the C-ABI in
include/cupic.h (libcupic.so, hand-written HIP for gfx950).  Names follow
the reference: ``init_sim``, ``charge_deposition``, ``poisson_solver``,
``field_solver``, ``particle_mover``, ``cc``, ``number_of_particles``
(src/init.h, src/mesh.h, src/particles.h, src/cc.h, src/diagnostic.h).
Errors raise :class:`CupicError` instead of exiting the process.

There is no CPU fallback: importing this package without the built HIP
library raises ImportError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcupic.so")

CUPIC_OK = 0
ERRORS = {-1: "ARG", -2: "HIP", -3: "ALLOC", -4: "CAPACITY", -5: "STATE", -6: "NUMERIC",
          -7: "IO"}

JACOBI_REFERENCE, JACOBI_RESIDUAL, JACOBI_REFINE, POISSON_DIRECT = 0, 1, 2, 3
DIRECT_AUTO, DIRECT_DENSE, DIRECT_FFT = 0, 1, 2
HINTS_OFF, HINTS_ON, HINTS_CHECKED = 0, 1, 2
QUIRK_Q7_REFERENCE_KICK = 1
QUIRK_Q10_SKIP_FIRST_LINE = 2
MESH_RHO, MESH_PHI, MESH_EX, MESH_EY, MESH_CHARGE = 0, 1, 2, 3, 4
ELECTRONS, IONS = 0, 1


class CupicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cupic error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
    """Mirror of cupic_params (include/cupic.h)."""
    _fields_ = [
        ("ncx", C.c_int32), ("ncy", C.c_int32),
        ("ds", C.c_double), ("dt", C.c_double), ("n", C.c_double),
        ("mi", C.c_double), ("kti", C.c_double), ("me", C.c_double), ("kte", C.c_double),
        ("phi_p", C.c_double), ("eps0", C.c_double),
        ("seed", C.c_uint64), ("run_id", C.c_uint64),
        ("quirks", C.c_uint32),
        ("jacobi_mode", C.c_int32), ("jacobi_min_sweeps", C.c_int32),
        ("jacobi_max_sweeps", C.c_int32), ("jacobi_check_every", C.c_int32),
        ("tile_cells", C.c_int32), ("direct_method", C.c_int32),
        ("jacobi_tol", C.c_double), ("jacobi_tol_init", C.c_double),
        ("Lx", C.c_double), ("Ly", C.c_double),
        ("dtin_e", C.c_double), ("dtin_i", C.c_double),
        ("per_row", C.c_int64),
    ]

    def copy(self):
        p = Params()
        C.pointer(p)[0] = self
        return p


class ServerConfig(C.Structure):
    """Mirror of cupic_server_config (include/cupic.h)."""
    _fields_ = [
        ("hbm_budget", C.c_int64), ("max_active", C.c_int32), ("round_steps", C.c_int32),
        ("window_steps", C.c_int32), ("max_steps", C.c_int32), ("rel_change", C.c_double),
        ("use_graph", C.c_int32), ("jacobi_mode", C.c_int32),
        ("jacobi_max_sweeps", C.c_int32), ("jacobi_check_every", C.c_int32),
        ("jacobi_tol", C.c_double), ("max_particles", C.c_int64), ("policy", C.c_int32),
    ]


class Point(C.Structure):
    """Mirror of cupic_point (include/cupic.h)."""
    _fields_ = [
        ("id", C.c_int64), ("status", C.c_int32), ("error", C.c_int32),
        ("phi_p", C.c_double), ("footprint", C.c_int64),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("steps", C.c_int64), ("windows", C.c_int64),
        ("current", C.c_double * 2), ("rel_change", C.c_double),
        ("total_probe", C.c_int64 * 2), ("particles", C.c_int64 * 2),
    ]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k in ("current", "total_probe", "particles")
                    else getattr(self, k)) for k, _ in self._fields_}


POINT_QUEUED, POINT_RUNNING, POINT_CONVERGED, POINT_CAPPED, POINT_FAILED = 0, 1, 2, 3, 4
SERVE_FCFS, SERVE_SMALLEST_FIRST = 0, 1


def _is_product_library(path):
    """A mapped shared object that carries this library's C-ABI: its name
    starts with libcupic and its dynamic string table holds cupic_init_dev
    (other libraries of the repository export no cupic_ symbols)."""
    base = os.path.basename(path)
    if not base.startswith("libcupic") or ".so" not in base:
        return False
    import mmap
    fname = path[:-len(" (deleted)")] if path.endswith(" (deleted)") else path
    try:
        with open(fname, "rb") as f, mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ) as m:
            return m.find(b"\0cupic_init_dev\0") >= 0
    except (OSError, ValueError):
        return True   # unreadable now (replaced or removed): count it as a copy


_OWN_ID = None   # (st_dev, st_ino) of the file this package mapped


def other_library_copies():
    """Copies of libcupic.so mapped into this process other than the one
    this package loaded (or, before loading, other than LIB_PATH): a file at
    another path, or another file at this path, read from /proc/self/maps."""
    own_id = _OWN_ID
    if own_id is None:
        try:
            own = os.stat(LIB_PATH)
            own_id = (own.st_dev, own.st_ino)
        except OSError:
            own_id = None
    others = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                parts = line.split(None, 5)
                if len(parts) < 6:
                    continue
                path = parts[5].strip()
                if not path.startswith("/") or not _is_product_library(path):
                    continue
                dev = tuple(int(x, 16) for x in parts[3].split(":"))
                ino = int(parts[4])
                # /proc reports the device as major:minor; compare the inode
                # and the path, since st_dev packs major/minor differently
                if own_id is not None and ino == own_id[1] and \
                        os.major(own_id[0]) == dev[0] and os.minor(own_id[0]) == dev[1]:
                    continue
                others.add(path)
    except OSError:
        pass
    return sorted(others)


class LibraryCopyError(ImportError):
    """A second copy of libcupic.so is mapped into the process."""


def _single_copy_check():
    """Refuses a process that holds a second copy of the library.  Two
    copies register the same kernel names, code objects and static state
    with the one HIP runtime of the process; that configuration is not
    supported (it is where a GPU memory fault was once seen)."""
    others = other_library_copies()
    if others:
        raise LibraryCopyError(f"another copy of libcupic is already mapped into this process "
                          f"({', '.join(others)}); load one copy per process")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python cupic_amd/build.py` "
                          "(there is no CPU fallback)")
    global _OWN_ID
    _single_copy_check()
    st = os.stat(LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    _OWN_ID = (st.st_dev, st.st_ino)
    P = C.POINTER(Params)
    vp = C.c_void_p
    i64p = C.POINTER(C.c_int64)
    i32p = C.POINTER(C.c_int32)
    dp = C.POINTER(C.c_double)
    sig = {
        "cupic_init_dev": [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int)],
        "cupic_params_default": [P],
        "cupic_params_derive": [P],
        "cupic_params_from_input_file": [C.c_char_p, P, i32p, i32p, i32p, i32p],
        "cupic_run_footprint": [P, C.c_int64, i64p],
        "cupic_run_create": [P, C.c_int64, C.POINTER(vp)],
        "cupic_run_destroy": [vp],
        "cupic_synchronize": [vp],
        "cupic_init_sim": [vp, dp],
        "cupic_init_sim_restart": [vp, C.c_int32, C.c_char_p, C.c_char_p, dp],
        "cupic_read_particle_file": [vp, C.c_int, C.c_char_p],
        "cupic_input_file_set_n_ini": [C.c_char_p, C.c_int32],
        "cupic_create_particles": [vp],
        "cupic_initialize_mesh": [vp],
        "cupic_adjust_leap_frog": [vp],
        "cupic_charge_deposition": [vp],
        "cupic_poisson_solver": [vp, C.c_double, i32p, dp],
        "cupic_poisson_solver_noalloc": [P, vp, vp, vp, C.c_double, vp, i32p, dp],
        "cupic_field_solver": [vp],
        "cupic_particle_mover": [vp],
        "cupic_cc": [vp, C.c_double],
        "cupic_step": [vp, C.c_double],
        "cupic_number_of_particles": [vp, C.c_int, i64p],
        "cupic_probe_tally": [vp, i64p, i64p, i64p, i64p, dp],
        "cupic_last_solve": [vp, i32p, dp],
        "cupic_solver_totals": [vp, i64p, i64p],
        "cupic_run_counters": [vp, i64p, i64p, i64p, i64p],
        "cupic_run_set_statistics": [vp, C.c_int32],
        "cupic_get_particles": [vp, C.c_int, vp, vp, vp, vp, C.c_int64, i64p],
        "cupic_set_particles": [vp, C.c_int, vp, vp, vp, vp, C.c_int64],
        "cupic_get_offsets": [vp, C.c_int, vp],
        "cupic_density_profile": [vp, C.c_int, vp],
        "cupic_get_bookmarks": [vp, C.c_int, vp],
        "cupic_get_mesh": [vp, C.c_int, vp, C.c_int64],
        "cupic_set_phi": [vp, vp, C.c_int64],
        "cupic_get_clock": [vp, dp, C.POINTER(C.c_uint64), i32p],
        "cupic_set_clock": [vp, dp, C.POINTER(C.c_uint64), C.c_int32],
        "cupic_particles_snapshot": [vp, C.c_int, C.c_char_p, C.c_double],
        "cupic_mesh_snapshot": [vp, C.c_int, C.c_char_p],
        "cupic_bookmarks_snapshot": [vp, C.c_int, C.c_char_p],
        "cupic_bins_snapshot": [vp, C.c_int, C.c_char_p],
        "cupic_batch_create": [C.POINTER(vp), C.c_int32, C.c_int32, C.POINTER(vp)],
        "cupic_batch_step": [vp, C.c_int32],
        "cupic_autotune_tile": [P, C.c_int32, i32p, dp, dp],
        "cupic_autotune_run": [vp, dp, C.c_int32, i32p, dp, dp],
        "cupic_run_set_tile": [vp, C.c_int32],
        "cupic_run_set_hint_mode": [vp, C.c_int32],
        "cupic_run_params": [vp, P],
        "cupic_batch_launch": [vp, C.c_int32],
        "cupic_batch_wait": [vp],
        "cupic_server_config_default": [C.POINTER(ServerConfig)],
        "cupic_server_create": [C.POINTER(ServerConfig), C.POINTER(vp)],
        "cupic_server_submit": [vp, P, i64p],
        "cupic_server_poll": [vp, i32p, i32p],
        "cupic_server_result": [vp, C.c_int64, C.POINTER(Point)],
        "cupic_server_stats": [vp, i64p, i64p, i64p, dp],
        "cupic_server_destroy": [vp],
        "cupic_batch_destroy": [vp],
        "cupic_timing_enable": [vp, C.c_int32],
        "cupic_timing_select": [vp, C.c_uint32],
        "cupic_timing_read": [vp, dp, i64p, C.c_int32],
        "cupic_bandwidth_probe": [C.c_int64, C.c_int32, C.c_int32, dp],
    }
    for name, args in sig.items():
        try:
            f = getattr(lib, name)
        except AttributeError as e:
            raise ImportError(f"{LIB_PATH} lacks {name}: rebuild with "
                              "`python cupic_amd/build.py`") from e
        f.argtypes = args
        f.restype = C.c_int
    lib.cupic_last_error.argtypes = [vp]
    lib.cupic_last_error.restype = C.c_char_p
    lib.cupic_batch_last_error.argtypes = [vp]
    lib.cupic_batch_last_error.restype = C.c_char_p
    lib.cupic_server_last_error.argtypes = [vp]
    lib.cupic_server_last_error.restype = C.c_char_p
    lib.cupic_timing_names.argtypes = []
    lib.cupic_timing_names.restype = C.c_char_p
    return lib


lib = _load()


def _check(code, handle=None, batch=False, server=False):
    if code != CUPIC_OK:
        fn = (lib.cupic_server_last_error if server else
              lib.cupic_batch_last_error if batch else lib.cupic_last_error)
        msg = fn(handle).decode(errors="replace")
        raise CupicError(code, msg)


def init_dev(device=0):
    """init_dev (src/init.cu:15-55): select a device; returns (name, CUs)."""
    buf = C.create_string_buffer(256)
    cu = C.c_int(0)
    _check(lib.cupic_init_dev(device, buf, 256, C.byref(cu)))
    return buf.value.decode(), cu.value


def params(**kw):
    """Config-1 defaults, overridden by keyword (``ppc`` sets n = ppc/ds^3)."""
    p = Params()
    _check(lib.cupic_params_default(C.byref(p)))
    ppc = kw.pop("ppc", None)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    if ppc is not None:
        p.n = ppc / (p.ds * p.ds * p.ds)
    p.eps0 = kw.get("eps0", 0.0)
    _check(lib.cupic_params_derive(C.byref(p)))
    return p


def input_file_set_n_ini(path, n_ini):
    """Sets line 2 of the reference's input file to "n_ini = <n_ini>;" (the
    rewrite at the end of src/main.cu:98-113)."""
    _check(lib.cupic_input_file_set_n_ini(str(path).encode(), n_ini))


def params_from_input_file(path):
    """Reads the reference's input/input_data (src/init.cu:424-789)."""
    p = Params()
    v = [C.c_int32() for _ in range(4)]
    _check(lib.cupic_params_from_input_file(path.encode(), C.byref(p), *[C.byref(x) for x in v]))
    return p, dict(n_ini=v[0].value, n_prev=v[1].value, n_save=v[2].value, n_fin=v[3].value)


def footprint(p, capacity=0):
    b = C.c_int64()
    _check(lib.cupic_run_footprint(C.byref(p), capacity, C.byref(b)))
    return b.value


def poisson_solver_noalloc(p, charge, phi, rho=None, tol=-1.0, stream=None):
    """The Jacobi solve of Run.poisson_solver (modes 0 and 1) on caller-owned
    device buffers (integer device addresses): charge int64 [ncy+1][ncx]
    read, phi FP32 [ncy+1][ncx] in/out, rho FP32 out or None.  Allocates no
    device memory: its temporaries live in a static device array of
    NOALLOC_SCRATCH_BYTES per GPU, shared by all calls on that GPU.  Returns
    (sweeps, residual) once the solve has finished."""
    s = C.c_int32()
    r = C.c_double()
    _check(lib.cupic_poisson_solver_noalloc(C.byref(p), charge, phi, rho, tol, stream,
                                            C.byref(s), C.byref(r)))
    return s.value, r.value


NOALLOC_SCRATCH_BYTES = 16 << 20


def bandwidth_probe(nbytes, kind="copy", reps=5):
    """Mean ms per launch of a plain streaming kernel over nbytes (copy:
    nbytes read and nbytes written; read; write) and the rate in bytes/s."""
    k = {"copy": 0, "read": 1, "write": 2}[kind]
    ms = C.c_double()
    _check(lib.cupic_bandwidth_probe(nbytes, k, reps, C.byref(ms)))
    moved = nbytes * (2 if k == 0 else 1)
    return ms.value, moved / (ms.value * 1e-3)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Run:
    """One simulation: owns its device buffers and stream (src/main.cu:42-44)."""

    def __init__(self, p, capacity=0):
        _single_copy_check()
        self.p = p.copy()
        self.h = C.c_void_p()
        _check(lib.cupic_run_create(C.byref(self.p), capacity, C.byref(self.h)))
        self.nnx = self.p.ncx + 1
        self.nny = self.p.ncy + 1

    def close(self):
        if self.h:
            lib.cupic_run_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, code):
        _check(code, self.h)

    # init (src/init.h:32-38)
    def init_sim(self):
        t = C.c_double()
        self._c(lib.cupic_init_sim(self.h, C.byref(t)))
        return t.value

    def init_sim_restart(self, n_ini, ions_path, electrons_path):
        """Resumes at t = n_ini dt from two particle files (src/init.cu:84-95);
        returns t."""
        t = C.c_double()
        self._c(lib.cupic_init_sim_restart(self.h, n_ini, str(ions_path).encode(),
                                           str(electrons_path).encode(), C.byref(t)))
        return t.value

    def read_particle_file(self, species, path):
        """Loads one species from a file in the snapshot text format."""
        self._c(lib.cupic_read_particle_file(self.h, species, str(path).encode()))

    def create_particles(self):
        self._c(lib.cupic_create_particles(self.h))

    def initialize_mesh(self):
        self._c(lib.cupic_initialize_mesh(self.h))

    def adjust_leap_frog(self):
        self._c(lib.cupic_adjust_leap_frog(self.h))

    # mesh (src/mesh.h:27-29)
    def charge_deposition(self):
        self._c(lib.cupic_charge_deposition(self.h))

    def poisson_solver(self, tol=-1.0):
        s = C.c_int32()
        r = C.c_double()
        self._c(lib.cupic_poisson_solver(self.h, tol, C.byref(s), C.byref(r)))
        return s.value, r.value

    def field_solver(self):
        self._c(lib.cupic_field_solver(self.h))

    # particles (src/particles.h:28) and cc (src/cc.h:29)
    def particle_mover(self):
        self._c(lib.cupic_particle_mover(self.h))

    def cc(self, t):
        self._c(lib.cupic_cc(self.h, t))

    def step(self, t):
        self._c(lib.cupic_step(self.h, t))

    def synchronize(self):
        self._c(lib.cupic_synchronize(self.h))

    # diagnostics (src/diagnostic.h:27-31)
    def number_of_particles(self, species):
        n = C.c_int64()
        self._c(lib.cupic_number_of_particles(self.h, species, C.byref(n)))
        return n.value

    def probe_tally(self):
        a = [(C.c_int64 * 2)() for _ in range(4)]
        q = C.c_double()
        self._c(lib.cupic_probe_tally(self.h, *a, C.byref(q)))
        last_probe, last_top, last_inj, tot = [np.array(list(x), np.int64) for x in a]
        return dict(probe=last_probe, top=last_top, injected=last_inj, total_probe=tot,
                    charge=q.value)

    def last_solve(self):
        s = C.c_int32()
        r = C.c_double()
        self._c(lib.cupic_last_solve(self.h, C.byref(s), C.byref(r)))
        return s.value, r.value

    def solver_totals(self):
        a, b = C.c_int64(), C.c_int64()
        self._c(lib.cupic_solver_totals(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def counters(self):
        """Since init_sim: far movers and fully pushed (undecided-hint)
        particles per species, steps, and particles moved summed over steps."""
        far, und = (C.c_int64 * 2)(), (C.c_int64 * 2)()
        n, ps = C.c_int64(), C.c_int64()
        self._c(lib.cupic_run_counters(self.h, far, und, C.byref(n), C.byref(ps)))
        return dict(far=[int(v) for v in far], undecided=[int(v) for v in und], steps=n.value,
                    particle_steps=ps.value)

    def set_statistics(self, on):
        """Counts fully pushed (undecided-hint) particles while on; speed only."""
        self._c(lib.cupic_run_set_statistics(self.h, 1 if on else 0))

    def particles(self, species):
        n = self.number_of_particles(species)
        a = [np.empty(n, np.float32) for _ in range(4)]
        m = C.c_int64()
        self._c(lib.cupic_get_particles(self.h, species, *[x.ctypes.data for x in a], n,
                                        C.byref(m)))
        return a

    def set_particles(self, species, x, y, vx, vy):
        a = [_f32(v) for v in (x, y, vx, vy)]
        self._c(lib.cupic_set_particles(self.h, species, *[v.ctypes.data for v in a], len(a[0])))

    def offsets(self, species):
        o = np.empty(self.p.ncx * self.p.ncy + 1, np.int32)
        self._c(lib.cupic_get_offsets(self.h, species, o.ctypes.data))
        return o

    def density_profile(self, species):
        """particles per cell row, shape (ncy,)"""
        a = np.empty(self.p.ncy, np.int64)
        self._c(lib.cupic_density_profile(self.h, species, a.ctypes.data))
        return a

    def bookmarks(self, species):
        """row bookmarks (first, last) per cell row, shape (ncy, 2); -1, -1 for an empty row"""
        a = np.empty(2 * self.p.ncy, np.int32)
        self._c(lib.cupic_get_bookmarks(self.h, species, a.ctypes.data))
        return a.reshape(self.p.ncy, 2)

    def mesh(self, which):
        """A mesh quantity in the reference layout, shape (nny, nnx)."""
        dt = {MESH_RHO: np.float32, MESH_PHI: np.float64, MESH_EX: np.float32,
              MESH_EY: np.float32, MESH_CHARGE: np.int64}[which]
        a = np.empty(self.nnx * self.nny, dt)
        self._c(lib.cupic_get_mesh(self.h, which, a.ctypes.data, a.size))
        return a.reshape(self.nny, self.nnx)

    def set_phi(self, phi):
        a = np.ascontiguousarray(phi, np.float64).ravel()
        self._c(lib.cupic_set_phi(self.h, a.ctypes.data, a.size))

    def clock(self):
        tin = (C.c_double * 2)()
        ninj = (C.c_uint64 * 2)()
        st = C.c_int32()
        self._c(lib.cupic_get_clock(self.h, tin, ninj, C.byref(st)))
        return list(tin), list(ninj), st.value

    def set_clock(self, tin, ninj, started=1):
        t = (C.c_double * 2)(*tin)
        n = (C.c_uint64 * 2)(*ninj)
        self._c(lib.cupic_set_clock(self.h, t, n, started))

    def particles_snapshot(self, species, path, t):
        self._c(lib.cupic_particles_snapshot(self.h, species, path.encode(), t))

    def mesh_snapshot(self, which, path):
        self._c(lib.cupic_mesh_snapshot(self.h, which, path.encode()))

    def bookmarks_snapshot(self, species, path):
        self._c(lib.cupic_bookmarks_snapshot(self.h, species, path.encode()))

    def bins_snapshot(self, species, path):
        self._c(lib.cupic_bins_snapshot(self.h, species, path.encode()))

    def set_tile(self, tile):
        """Cells per wave strip of the push passes (0: from ppc); speed only."""
        self._c(lib.cupic_run_set_tile(self.h, tile))
        self.p.tile_cells = tile

    def set_hint_mode(self, mode):
        """Push hints of the counting pass: HINTS_OFF, HINTS_ON (default) or
        HINTS_CHECKED (every hint verified against the push); speed only."""
        self._c(lib.cupic_run_set_hint_mode(self.h, mode))

    def autotune(self, t, steps=3):
        """Tunes the tile on this run by stepping it from time t; returns
        (next t, tile, ms per step, {tile: ms})."""
        tt = C.c_double(t)
        best = C.c_int32()
        ms = C.c_double()
        times = (C.c_double * 17)()
        self._c(lib.cupic_autotune_run(self.h, C.byref(tt), steps, C.byref(best), C.byref(ms),
                                       times))
        self.p.tile_cells = best.value
        return tt.value, best.value, ms.value, {c: times[c] for c in range(17) if times[c] > 0}

    def timing(self, on=True, classes=None):
        """Per-class device time with HIP events; `classes` (names from
        timing_read) limits the events to those classes."""
        if on and classes:
            names = lib.cupic_timing_names().decode().split(",")
            mask = 0
            for c in classes:
                mask |= 1 << names.index(c)
            self._c(lib.cupic_timing_select(self.h, mask))
        else:
            self._c(lib.cupic_timing_enable(self.h, 1 if on else 0))

    def timing_read(self):
        names = lib.cupic_timing_names().decode().split(",")
        ms = (C.c_double * len(names))()
        n = (C.c_int64 * len(names))()
        self._c(lib.cupic_timing_read(self.h, ms, n, len(names)))
        return {k: (ms[i], n[i]) for i, k in enumerate(names)}


class Batch:
    """Runs advanced together: one launch per kernel per step, optionally one
    HIP graph per step."""

    def __init__(self, runs, graph=True):
        self.runs = list(runs)
        arr = (C.c_void_p * len(self.runs))(*[r.h.value for r in self.runs])
        self.h = C.c_void_p()
        _check(lib.cupic_batch_create(arr, len(self.runs), 1 if graph else 0, C.byref(self.h)))

    def step(self, n=1):
        _check(lib.cupic_batch_step(self.h, n), self.h, batch=True)

    def launch(self, n=1):
        _check(lib.cupic_batch_launch(self.h, n), self.h, batch=True)

    def wait(self):
        _check(lib.cupic_batch_wait(self.h), self.h, batch=True)

    def close(self):
        if self.h:
            lib.cupic_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_tile(p, steps=5):
    """Fastest tile_cells for runs shaped like p (push passes timed with HIP
    events); returns (tile, ms per step, {tile: ms})."""
    best = C.c_int32()
    ms = C.c_double()
    times = (C.c_double * 17)()
    q = p.copy()
    _check(lib.cupic_autotune_tile(C.byref(q), steps, C.byref(best), C.byref(ms), times))
    return best.value, ms.value, {c: times[c] for c in range(17) if times[c] > 0}


def server_config(**kw):
    """Server defaults (include/cupic.h cupic_server_config_default), overridden
    by keyword."""
    c = ServerConfig()
    _check(lib.cupic_server_config_default(C.byref(c)))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Server:
    """Continuous batching of I-V sweep points (one request = one bias point):
    resident runs advance together in batched rounds until their probe current
    is steady; requests are admitted in arrival order (or fewest particles
    first) by slots, device-memory budget and resident-particle budget."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else server_config(**kw)
        self.h = C.c_void_p()
        _check(lib.cupic_server_create(C.byref(self.cfg), C.byref(self.h)))

    def submit(self, p):
        i = C.c_int64()
        _check(lib.cupic_server_submit(self.h, C.byref(p), C.byref(i)), self.h, server=True)
        return i.value

    def poll(self):
        """One round; returns (resident runs, queued requests)."""
        a, q = C.c_int32(), C.c_int32()
        _check(lib.cupic_server_poll(self.h, C.byref(a), C.byref(q)), self.h, server=True)
        return a.value, q.value

    def result(self, i):
        pt = Point()
        _check(lib.cupic_server_result(self.h, i, C.byref(pt)), self.h, server=True)
        return pt.as_dict()

    def stats(self):
        r, b, u = C.c_int64(), C.c_int64(), C.c_int64()
        t = C.c_double()
        _check(lib.cupic_server_stats(self.h, C.byref(r), C.byref(b), C.byref(u), C.byref(t)),
               self.h, server=True)
        return dict(rounds=r.value, rebuilds=b.value, hbm_used=u.value, clock=t.value)

    def close(self):
        if self.h:
            lib.cupic_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
