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
