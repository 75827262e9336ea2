"""Snapshots in the reference's text formats (src/diagnostic.cu:48-133), the
restart path (src/init.cu:84-95, 327-420) and the C++ host driver
(cupic_amd/cupic_main, the counterpart of src/main.cu) end to end: its
snapshot files equal those of the same run stepped through the Python
binding, and it resumes from its own snapshots."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

cp = pytest.importorskip("cupic_amd")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cupic_amd", "cupic_main")
LINE = re.compile(r"^ (-?\d\.\d{17}e[+-]\d\d) (-?\d\.\d{17}e[+-]\d\d) "
                  r"(-?\d\.\d{17}e[+-]\d\d) (-?\d\.\d{17}e[+-]\d\d) $")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    return cp.init_dev(0)


def test_particle_and_mesh_snapshots(dev, tmp_path):
    p = cp.params(ncx=24, ncy=10, ppc=6)
    run = cp.Run(p)
    run.init_sim()
    run.step(p.dt)
    base = str(tmp_path / "electrons_t_1")
    run.particles_snapshot(0, base, p.dt)
    lines = open(base + ".dat").read().split("\n")
    assert lines[-1] == ""
    vals = []
    for ln in lines[:-1]:
        m = LINE.match(ln)
        assert m, repr(ln)
        vals.append([float(v) for v in m.groups()])
    got = np.array(vals)
    want = np.stack([a.astype(np.float64) for a in run.particles(0)], axis=1)
    assert np.array_equal(got, want)
    mb = str(tmp_path / "potential_t_1")
    run.mesh_snapshot(cp.MESH_PHI, mb)
    blocks = open(mb + ".dat").read().split("\n\n")
    phi = run.mesh(cp.MESH_PHI)
    nny, nnx = phi.shape
    assert len(blocks) == nnx + 1 and blocks[-1] == ""
    for i, blk in enumerate(blocks[:-1]):
        rows = blk.strip("\n").split("\n")
        assert len(rows) == nny
        for j, ln in enumerate(rows):
            a, b, v = ln.split()
            assert ln.startswith(" ") and ln.endswith(" ")
            assert (int(a), int(b)) == (i, j) and float(v) == phi[j, i]
    run.close()


def _rows_of(y, ds):
    # int(y/ds) with the float y widened to double (src/diagnostic.cu:204)
    return (y.astype(np.float64) / ds).astype(np.int64)


@pytest.mark.parametrize("case", ["stepped", "empty_rows"])
def test_bookmarks_and_bins(dev, tmp_path, case):
    """save_bm / save_bins (src/diagnostic.cu:137-215): the bookmarks follow
    from the rows of the downloaded particles, first and last index per row and
    -1, -1 for an empty row (src/init.cu:401-407), and the files carry the
    reference's line formats."""
    p = cp.params(ncx=24, ncy=10, ppc=6)
    run = cp.Run(p)
    run.init_sim()
    if case == "stepped":
        t = p.dt
        for _ in range(5):
            run.step(t)
            t += p.dt
    else:
        # rows 1, 2, 6 and 9 left empty; the rest hold uneven counts
        rng = np.random.default_rng(7)
        rows = np.repeat([0, 3, 4, 5, 7, 8], [5, 1, 9, 2, 4, 3])
        n = rows.size
        x = (rng.random(n) * p.ncx * p.ds).astype(np.float32)
        y = ((rows + rng.uniform(0.05, 0.95, n)) * p.ds).astype(np.float32)
        v = rng.standard_normal((2, n)).astype(np.float32)
        run.set_particles(0, x, y, v[0], v[1])
    for s in (0, 1):
        bm = run.bookmarks(s)
        y = run.particles(s)[1]
        rows = _rows_of(y, p.ds)
        assert np.all(np.diff(rows) >= 0)
        want = np.full((p.ncy, 2), -1, np.int32)
        for j in range(p.ncy):
            idx = np.nonzero(rows == j)[0]
            if idx.size:
                want[j] = idx[0], idx[-1]
        assert np.array_equal(bm, want)
        if case == "empty_rows" and s == 0:
            assert np.array_equal(np.nonzero(bm[:, 0] < 0)[0], [1, 2, 6, 9])
        assert np.array_equal(np.where(bm[:, 0] >= 0, bm[:, 1] - bm[:, 0] + 1, 0),
                              run.density_profile(s))
        base = str(tmp_path / f"bm_{s}")
        run.bookmarks_snapshot(s, base)
        lines = open(base + ".dat").read().split("\n")
        assert lines[-1] == "" and len(lines) == p.ncy + 1
        for j, ln in enumerate(lines[:-1]):
            assert ln == f" {2 * j} {bm[j, 0]} {bm[j, 1]} "
        base = str(tmp_path / f"bins_{s}")
        run.bins_snapshot(s, base)
        lines = open(base + ".dat").read().split("\n")
        assert lines[-1] == "" and len(lines) == y.size + 1
        for k, ln in enumerate(lines[:-1]):
            assert ln == f" {k} {rows[k]} "
    run.close()


def test_host_driver_matches_binding(dev, tmp_path):
    assert os.path.exists(DRIVER), "cupic_main is built by cupic_amd/build.py"
    say("cupic_main --config small --steps 20")
    r = subprocess.run([DRIVER, "--config", "small", "--steps", "20", "--out", str(tmp_path),
                        "--save-every", "10", "--quiet"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation finished!" in r.stdout
    for i in (0, 10, 20):
        for sp in ("electrons", "ions"):
            assert (tmp_path / f"{sp}_t_{i}.dat").exists()
    # the same run through the binding, t accumulated as src/main.cu:57,59
    p = cp.params()
    run = cp.Run(p)
    run.init_sim()
    t = 0.0
    t += p.dt
    for i in range(1, 21):
        run.step(t)
        t += p.dt
    ref = str(tmp_path / "binding_electrons_t_20")
    run.particles_snapshot(0, ref, t)
    assert open(ref + ".dat").read() == open(tmp_path / "electrons_t_20.dat").read()
    run.close()


def same_particles(a, b, s):
    for u, v in zip(a.particles(s), b.particles(s)):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))


def test_particle_file_round_trip(dev, tmp_path):
    """A snapshot read back holds the same FP32 state in the same order
    (%.17e round-trips every float); with quirk Q10 the reader drops the first
    line as the reference's does (src/init.cu:372,392); a malformed line is an
    I/O error."""
    p = cp.params(ncx=24, ncy=10, ppc=6)
    a = cp.Run(p)
    a.init_sim()
    t = p.dt
    for _ in range(4):
        a.step(t)
        t += p.dt
    for s, name in ((0, "electrons"), (1, "ions")):
        a.particles_snapshot(s, str(tmp_path / name), t)
    b = cp.Run(p)
    b.read_particle_file(0, tmp_path / "electrons.dat")
    b.read_particle_file(1, tmp_path / "ions.dat")
    for s in (0, 1):
        same_particles(a, b, s)
    q = cp.params(ncx=24, ncy=10, ppc=6, quirks=cp.QUIRK_Q10_SKIP_FIRST_LINE)
    c = cp.Run(q)
    c.read_particle_file(0, tmp_path / "electrons.dat")
    got = c.particles(0)
    want = a.particles(0)
    assert len(got[0]) == len(want[0]) - 1
    for u, v in zip(got, want):
        assert np.array_equal(u.view(np.uint32), v[1:].view(np.uint32))
    bad = tmp_path / "bad.dat"
    bad.write_text(" 1.0 2.0 3.0 4.0 \n 1.0 2.0 x \n")
    with pytest.raises(cp.CupicError) as e:
        b.read_particle_file(0, bad)
    assert e.value.code == -7
    for r in (a, b, c):
        r.close()


def test_restart_from_snapshots(dev, tmp_path):
    """init_sim with n_ini > 0: t = n_ini dt, the particles of the files, a
    fresh solve from the linear guess; the run then steps on."""
    p = cp.params(ncx=32, ncy=16, ppc=10)
    a = cp.Run(p)
    a.init_sim()
    t = 0.0
    t += p.dt
    for i in range(1, 7):
        a.step(t)
        t += p.dt
    t_save = t - p.dt  # the time of step 6, as the driver passes it
    a.particles_snapshot(0, str(tmp_path / "electrons"), t_save)
    a.particles_snapshot(1, str(tmp_path / "ions"), t_save)
    b = cp.Run(p)
    t_b = b.init_sim_restart(6, tmp_path / "ions.dat", tmp_path / "electrons.dat")
    assert t_b == 6 * p.dt
    for s in (0, 1):
        same_particles(a, b, s)
    sweeps, resid = b.last_solve()
    assert sweeps > 0
    t_b += p.dt
    for _ in range(5):
        b.step(t_b)
        t_b += p.dt
    tally = b.probe_tally()
    assert tally["injected"].sum() > 0
    with pytest.raises(cp.CupicError):
        b.init_sim_restart(0, tmp_path / "ions.dat", tmp_path / "electrons.dat")
    a.close()
    b.close()


def test_host_driver_restart_and_input_rewrite(dev, tmp_path):
    """cupic_main resumes from its own snapshots (copied to ions.dat and
    electrons.dat, as the reference expects them) and, given an input file,
    sets its n_ini line to the last step."""
    assert os.path.exists(DRIVER), "cupic_main is built by cupic_amd/build.py"
    out1, rdir, out2 = tmp_path / "o1", tmp_path / "r", tmp_path / "o2"
    for d in (out1, rdir, out2):
        d.mkdir()
    say("cupic_main --config small --steps 10, then --restart 10 --steps 14")
    r = subprocess.run([DRIVER, "--config", "small", "--steps", "10", "--out", str(out1),
                        "--save-every", "10", "--quiet"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for sp in ("electrons", "ions"):
        (rdir / f"{sp}.dat").write_text((out1 / f"{sp}_t_10.dat").read_text())
    r = subprocess.run([DRIVER, "--config", "small", "--restart", "10", "--restart-dir",
                        str(rdir), "--steps", "14", "--out", str(out2), "--save-every", "2",
                        "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation state loaded from time t = 0.5" in r.stdout
    assert "Simulation finished!" in r.stdout
    # the snapshot written at the start of the resumed run is the state loaded
    for sp in ("electrons", "ions"):
        assert (out2 / f"{sp}_t_10.dat").read_text() == (out1 / f"{sp}_t_10.dat").read_text()
        assert (out2 / f"{sp}_t_14.dat").exists()
    # an input file of the reference's format, run for 3 steps: n_ini becomes 3
    inp = tmp_path / "input_data"
    inp.write_text("#execution configuration\nn_ini = 0;\nn_prev = 0;\nn_save = 100;\n"
                   "n_fin = 200000;\nne = 1.0e14;\nTe = 1.0e3;\nbeta = 1.0e-1;\n"
                   "gamma = 1.0e3;\nphi_p = -5.0e0;\nncx = 16;\nncy = 8;\nds = 1.0e-1;\n"
                   "dt = 1.0e-3;\n")
    r = subprocess.run([DRIVER, "--input", str(inp), "--steps", "3", "--quiet"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert inp.read_text().split("\n")[1] == "n_ini = 3;"
