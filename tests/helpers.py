"""Shared helpers for the GPU parity tests: move state between the HIP runs
(cupic_amd, through the C-ABI) and the CPU oracle."""
import ctypes as C

import numpy as np

from oracle import oracle as O


def oracle_params(p):
    """Oracle parameters equal to a cupic_amd Params."""
    o = O.Params()
    for f in ("ncx", "ncy", "ds", "dt", "n", "mi", "kti", "me", "kte", "phi_p", "seed",
              "run_id", "quirks"):
        setattr(o, f, getattr(p, f))
    o.eps0 = 0.0
    O.lib().orc_params_derive(C.byref(o))
    assert o.per_row == p.per_row and o.dtin_e == p.dtin_e and o.eps0 == p.eps0
    return o


def interleaved_E(run):
    """GPU field as the oracle's interleaved (Ex, Ey) [ncy+1][ncx] float32."""
    import cupic_amd as cp
    ex = periodic(run.mesh(cp.MESH_EX))
    ey = periodic(run.mesh(cp.MESH_EY))
    e = np.empty(ex.size * 2, np.float32)
    e[0::2] = ex.ravel()
    e[1::2] = ey.ravel()
    return e


def periodic(a):
    """Reference layout (nny, nnx) -> periodic (nny, ncx), checking the image column."""
    assert np.array_equal(a[:, -1], a[:, 0])
    return np.ascontiguousarray(a[:, :-1])


def mirror_init(op, s):
    N = op.per_row * op.ncy
    m = [np.empty(N, np.float32) for _ in range(4)]
    off = np.zeros(op.ncx * op.ncy + 1, np.int64)
    O.lib().orc_mirror_init(op, s, *m, off)
    return m, off


def injection_for_step(op, s, t, tin, started, ninj, E):
    """Oracle injection of species s at cc(t) from the run's clock state."""
    dtin = op.dtin_e if s == 0 else op.dtin_i
    tin_c = C.c_double(tin if started else t + dtin)
    seq = np.zeros(1 << 16)
    n_in = O.lib().orc_injection_count(t, op.dt, C.byref(tin_c), dtin, seq.ctypes.data,
                                       len(seq))
    inj = [np.empty(n_in, np.float32) for _ in range(4)]
    O.lib().orc_inject_f32(op, s, int(ninj), n_in, t, seq[:max(n_in, 1)].copy(), E, *inj)
    return inj, tin_c.value


def mirror_step(op, s, pre, E, inj):
    c = O.f32c(op)
    n = len(pre[0])
    n_in = len(inj[0])
    out = [np.empty(n + n_in, np.float32) for _ in range(4)]
    off = np.zeros(op.ncx * op.ncy + 1, np.int64)
    pr, tp, nf = C.c_int64(), C.c_int64(), C.c_int64()
    kept = O.lib().orc_mirror_step(c, s, n, *pre, E, n_in, *inj, *out, off, C.byref(pr),
                                   C.byref(tp), C.byref(nf))
    return [o[:kept] for o in out], off, pr.value, tp.value, nf.value


def fixed_charge(op, states):
    """int64 fixed-point node charge of electrons (states[0]) and ions (states[1])."""
    c = O.f32c(op)
    acc = np.zeros(op.ncx * (op.ncy + 1), np.int64)
    for s, st in enumerate(states):
        O.lib().orc_deposit_fixed_f32(len(st[0]), st[0], st[1], c, -1 if s == 0 else 1, acc)
    return acc.reshape(op.ncy + 1, op.ncx)


def mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32)) if a.dtype == np.float32 \
        else np.flatnonzero(a != b)
    return f"{len(bad)} mismatches, first {bad[:5]}"


def rel_err(a, b):
    """Norm-wise relative error max|a-b| / max|b|."""
    b = np.asarray(b, np.float64)
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


class MirrorRun:
    """A whole run replayed by the oracle's FP32 mirror, independently of the
    GPU: initial load and sort, fixed-point deposit, the mode-1 FP32 Jacobi
    with FP64 residual, the field, the half-step stagger, then per step the
    cycle deposit -> solve -> field -> push/absorb/inject/sort (the order of
    init_sim and of the time loop, src/init.cu:57-102, src/main.cu:61-73)."""

    def __init__(self, p):
        self.op = op = oracle_params(p)
        self.c = O.f32c(op)
        self.tol, self.tol_init = p.jacobi_tol, p.jacobi_tol_init
        self.max, self.check = p.jacobi_max_sweeps, p.jacobi_check_every
        self.nn = op.ncx * (op.ncy + 1)
        phi = np.zeros(self.nn)
        O.lib().orc_phi_linear(op.ncx, op.ncy, op.phi_p, phi)
        self.phi = phi.astype(np.float32)
        self.st = [list(mirror_init(op, s)[0]) for s in (0, 1)]
        self.off = [None, None]
        self.tin, self.ninj, self.started = [0.0, 0.0], [0, 0], False
        self.total_probe = [0, 0]
        self.sweeps, self.resid = self.solve(self.tol_init)
        self.E = self.field()
        for s in (0, 1):
            x, y, vx, vy = self.st[s]
            O.lib().orc_mirror_half_kick(self.c, s, len(x), x, y, vx, vy, self.E)

    def solve(self, tol):
        acc = fixed_charge(self.op, self.st).ravel()
        b64 = np.empty(self.nn)
        b32 = np.empty(self.nn, np.float32)
        O.lib().orc_mirror_rhs(self.op, acc, b64, b32)
        res = C.c_double()
        sw = O.lib().orc_mirror_solve_f32(self.op.ncx, self.op.ncy, b64, b32, self.phi, tol,
                                          self.max, self.check, C.byref(res))
        return sw, res.value

    def field(self):
        ex = np.empty(self.nn)
        ey = np.empty(self.nn)
        O.lib().orc_field(self.op.ncx, self.op.ncy, self.op.ds, self.phi.astype(np.float64),
                          ex, ey)
        E = np.empty(2 * self.nn, np.float32)
        E[0::2] = ex.astype(np.float32)
        E[1::2] = ey.astype(np.float32)
        return E

    def step(self, t):
        self.sweeps, self.resid = self.solve(self.tol)
        self.E = self.field()
        for s in (0, 1):
            inj, tin_after = injection_for_step(self.op, s, t, self.tin[s], self.started,
                                                self.ninj[s], self.E)
            out, off, pr, tp, nf = mirror_step(self.op, s, self.st[s], self.E, inj)
            self.st[s] = out
            self.off[s] = off
            self.total_probe[s] += pr
            self.ninj[s] += len(inj[0])
            self.tin[s] = tin_after
        self.started = True

    def row_profile(self, s):
        """particles per cell row"""
        rows = np.floor(self.st[s][1].astype(np.float64) / self.op.ds).astype(np.int64)
        return np.bincount(np.minimum(rows, self.op.ncy - 1), minlength=self.op.ncy)
