"""Writes the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

The reference cannot be built or run here (CUDA-only, needs GSL) and its
own tests hold no vectors, so the fixtures are produced by the oracle,
whose functions are pinned by closed forms, invariants and library routines
in tests/test_oracle_*.py.  They freeze, per case:

  init_s      initial load of species s after the cell sort: (4, N) float32
              x, y, vx, vy, and off_s, the cell offsets (src/init.cu:106-198)
  charge0     int64 fixed-point node charge of the initial load, [ncy+1][ncx]
              (2^-32 e per unit; src/mesh.cu:162-293)
  phi         FP64 converged potential of that charge with the Dirichlet
              rows (src/mesh.cu:55-120, src/init.cu:235-241), [ncy+1][ncx]
  phi32, E    phi rounded to float32, and E = -grad phi evaluated from it
              (src/mesh.cu:385-472), interleaved (Ex, Ey) float32
  kick_s      (vx, vy) after the half-step stagger with E (src/init.cu:793-833)
  step_s      state after one push + absorb + inject + sort at t = dt
              (src/particles.cu:110-231, src/cc.cu:134-296), off1_s its
              offsets, counts_s = (probe, top, injected), tin_s the clock
  charge1     fixed-point node charge of the stepped state

Grids are periodic in x ([ncy+1][ncx]; the reference's column nnx-1 is the
image of column 0).
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SEED = 20261203

CASES = {
    "small_32x16": dict(ncx=32, ncy=16, ppc=8, mi=100.0, kti=1.0, phi_p=-5.0),
    "ragged_37x9": dict(ncx=37, ncy=9, ppc=7, mi=100.0, kti=1.0, phi_p=2.0),
}


def fixed_charge(op, states):
    c = O.f32c(op)
    acc = np.zeros(op.ncx * (op.ncy + 1), np.int64)
    for s, st in enumerate(states):
        O.lib().orc_deposit_fixed_f32(len(st[0]), st[0], st[1], c, -1 if s == 0 else 1, acc)
    return acc.reshape(op.ncy + 1, op.ncx)


def converged_phi(op, charge):
    """FP64 Jacobi from the linear start until the update is below 1e-13."""
    ncx, ncy = op.ncx, op.ncy
    q = charge.astype(np.float64).ravel() * 2.0 ** -32
    rho = np.zeros_like(q)
    O.lib().orc_rho_from_Q(ncx, ncy, op.ds, q, rho)
    phi = np.zeros_like(q)
    O.lib().orc_phi_linear(ncx, ncy, op.phi_p, phi)
    res = C.c_double()
    O.lib().orc_poisson_solve(ncx, ncy, op.ds, op.eps0, rho, phi, 1, 1e-13, 0, 400000, 100,
                              C.byref(res))
    assert res.value < 1e-13, res.value
    return phi.reshape(ncy + 1, ncx)


def make_case(cfg):
    op = O.make_params(seed=SEED, **cfg)
    out = {"ncx": op.ncx, "ncy": op.ncy, "ds": op.ds, "dt": op.dt, "n": op.n, "mi": op.mi,
           "kti": op.kti, "phi_p": op.phi_p, "seed": op.seed, "run_id": op.run_id,
           "per_row": op.per_row}
    init = []
    for s in (0, 1):
        N = op.per_row * op.ncy
        m = [np.empty(N, np.float32) for _ in range(4)]
        off = np.zeros(op.ncx * op.ncy + 1, np.int64)
        O.lib().orc_mirror_init(op, s, *m, off)
        init.append(m)
        out[f"init_{s}"] = np.stack(m)
        out[f"off_{s}"] = off.astype(np.int32)
    out["charge0"] = fixed_charge(op, init)
    phi = converged_phi(op, out["charge0"])
    out["phi"] = phi
    phi32 = phi.astype(np.float32)
    out["phi32"] = phi32
    ex = np.zeros(phi.size)
    ey = np.zeros(phi.size)
    O.lib().orc_field(op.ncx, op.ncy, op.ds, phi32.astype(np.float64).ravel(), ex, ey)
    E = np.empty(2 * phi.size, np.float32)
    E[0::2] = ex.astype(np.float32)
    E[1::2] = ey.astype(np.float32)
    out["E"] = E
    c = O.f32c(op)
    t = op.dt
    posts = []
    for s in (0, 1):
        x, y, vx, vy = (a.copy() for a in init[s])
        O.lib().orc_mirror_half_kick(c, s, len(x), x, y, vx, vy, E)
        out[f"kick_{s}"] = np.stack([vx, vy])
        # injection of the first cc: the clock starts at t + dtin (src/cc.cu:35-36)
        dtin = op.dtin_e if s == 0 else op.dtin_i
        tin = C.c_double(t + dtin)
        seq = np.zeros(1 << 16)
        n_in = O.lib().orc_injection_count(t, op.dt, C.byref(tin), dtin, seq.ctypes.data,
                                           len(seq))
        inj = [np.empty(n_in, np.float32) for _ in range(4)]
        O.lib().orc_inject_f32(op, s, 0, n_in, t, seq[:max(n_in, 1)].copy(), E, *inj)
        n = len(x)
        res = [np.empty(n + n_in, np.float32) for _ in range(4)]
        off = np.zeros(op.ncx * op.ncy + 1, np.int64)
        pr, tp, nf = C.c_int64(), C.c_int64(), C.c_int64()
        kept = O.lib().orc_mirror_step(c, s, n, x, y, vx, vy, E, n_in, *inj, *res, off,
                                       C.byref(pr), C.byref(tp), C.byref(nf))
        res = [r[:kept] for r in res]
        posts.append(res)
        out[f"step_{s}"] = np.stack(res)
        out[f"off1_{s}"] = off.astype(np.int32)
        out[f"counts_{s}"] = np.array([pr.value, tp.value, n_in], np.int64)
        out[f"tin_{s}"] = np.float64(tin.value)
    out["charge1"] = fixed_charge(op, posts)
    return out


def main():
    for name, cfg in CASES.items():
        d = make_case(cfg)
        path = os.path.join(HERE, f"golden_{name}.npz")
        np.savez_compressed(path, **d)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
