"""Golden fixtures (tests/golden/*.npz): the oracle still writes them bit for
bit, and each fixture satisfies what the mathematics fixes independently of
the oracle's code (charge and count conservation, sortedness, the discrete
Poisson equation, a direct sparse solve, central differences)."""
import glob
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

FILES = sorted(glob.glob(os.path.join(HERE, "golden", "golden_*.npz")))
CASES = {os.path.basename(f)[7:-4]: f for f in FILES}


def load(name):
    with np.load(CASES[name], allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert set(CASES) == set(make_golden.CASES)


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_oracle_reproduces_fixture(name):
    want = load(name)
    got = make_golden.make_case(make_golden.CASES[name])
    assert set(got) == set(want)
    for k, v in want.items():
        g = np.atleast_1d(np.asarray(got[k], dtype=v.dtype))
        v = np.atleast_1d(v)
        assert g.shape == v.shape, k
        assert g.tobytes() == v.tobytes(), k


def cell_keys(x, y, ds, ncx, ncy):
    ic = np.clip(np.floor(x.astype(np.float64) / ds).astype(np.int64), 0, ncx - 1)
    jc = np.clip(np.floor(y.astype(np.float64) / ds).astype(np.int64), 0, ncy - 1)
    return jc * ncx + ic


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_fixture_invariants(name):
    d = load(name)
    ncx, ncy, ds = int(d["ncx"]), int(d["ncy"]), float(d["ds"])
    Lx, Ly = ncx * ds, ncy * ds
    N = int(d["per_row"]) * ncy
    for tag, off_tag in (("init", "off"), ("step", "off1")):
        for s in (0, 1):
            st = d[f"{tag}_{s}"]
            x, y = st[0], st[1]
            assert np.all((x >= 0) & (x < Lx) & (y >= 0) & (y < Ly))
            keys = cell_keys(x, y, ds, ncx, ncy)
            assert np.all(np.diff(keys) >= 0), "sorted by cell key"
            off = d[f"{off_tag}_{s}"].astype(np.int64)
            want = np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=ncx * ncy))])
            assert np.array_equal(off, want)
    # initial load: per_row particles in every row, y inside its row
    for s in (0, 1):
        y = d[f"init_{s}"][1]
        assert len(y) == N
        rows = np.floor(y.astype(np.float64) / ds).astype(np.int64)
        assert np.array_equal(np.bincount(rows, minlength=ncy), np.full(ncy, d["per_row"]))
    # count conservation N' = N - probe - top + injected
    for s in (0, 1):
        pr, tp, inj = d[f"counts_{s}"]
        assert d[f"step_{s}"].shape[1] == N - pr - tp + inj
    # exact fixed-point charge: 2^32 per particle, electrons -1, ions +1
    assert d["charge0"].sum() == 0
    n_e, n_i = d["step_0"].shape[1], d["step_1"].shape[1]
    assert d["charge1"].sum() == (n_i - n_e) * (1 << 32)


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_fixture_potential_solves_poisson(name):
    """phi satisfies the 5-point system with the Dirichlet rows, and equals a
    direct sparse solve of that system (scipy)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    d = load(name)
    ncx, ncy, ds, n = int(d["ncx"]), int(d["ncy"]), float(d["ds"]), float(d["n"])
    phi = d["phi"]
    assert np.allclose(phi[0], d["phi_p"], rtol=0, atol=0) and np.all(phi[ncy] == 0.0)
    q = d["charge0"].astype(np.float64) * 2.0 ** -32
    rho = q / ds ** 3                 # interior rows
    b = ds * ds * rho / n             # eps0 = n in normalised units
    inner = phi[1:ncy]
    nb = np.roll(phi, 1, 1) + np.roll(phi, -1, 1)
    res = 0.25 * (b[1:ncy] + nb[1:ncy] + phi[0:ncy - 1] + phi[2:ncy + 1]) - inner
    assert np.abs(res).max() < 1e-12
    # direct solve of 4 phi - neighbours = b on rows 1..ncy-1, periodic in x
    m = (ncy - 1) * ncx
    idx = np.arange(m).reshape(ncy - 1, ncx)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.full(m, 4.0)]
    for di in (-1, 1):
        rows.append(idx.ravel())
        cols.append(np.roll(idx, di, 1).ravel())
        vals.append(np.full(m, -1.0))
    rows.append(idx[1:].ravel()); cols.append(idx[:-1].ravel()); vals.append(np.full(m - ncx, -1.0))
    rows.append(idx[:-1].ravel()); cols.append(idx[1:].ravel()); vals.append(np.full(m - ncx, -1.0))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(m, m))
    rhs = b[1:ncy].copy()
    rhs[0] += phi[0]
    rhs[-1] += phi[ncy]
    direct = spl.spsolve(A.tocsc(), rhs.ravel()).reshape(ncy - 1, ncx)
    assert np.abs(direct - inner).max() < 1e-9 * max(1.0, np.abs(inner).max())


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_fixture_field_is_central_difference(name):
    d = load(name)
    ncy, ds = int(d["ncy"]), float(d["ds"])
    p = d["phi32"].astype(np.float64)
    E = d["E"]
    ex = E[0::2].reshape(p.shape)
    ey = E[1::2].reshape(p.shape)
    want_ex = (np.roll(p, 1, 1) - np.roll(p, -1, 1)) / (2 * ds)
    want_ey = np.empty_like(p)
    want_ey[1:ncy] = (p[0:ncy - 1] - p[2:ncy + 1]) / (2 * ds)
    want_ey[0] = (p[0] - p[1]) / ds
    want_ey[ncy] = (p[ncy - 1] - p[ncy]) / ds
    want_ex[0] = 0.0
    want_ex[ncy] = 0.0
    assert np.array_equal(ex, want_ex.astype(np.float32))
    assert np.array_equal(ey, want_ey.astype(np.float32))
