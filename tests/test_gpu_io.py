"""Snapshots in the reference's text formats (src/diagnostic.cu:48-133) and
the C++ host driver (cupic_amd/cupic_main, the counterpart of src/main.cu)
end to end: its snapshot files equal those of the same run stepped through
the Python binding."""
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
