"""CPU checks of the C-ABI library: it loads, exports every function that
include/cupic.h declares, and its host-side logic (parameter derivation,
input-file reader, footprint) agrees with the oracle.  No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cupic.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cupic_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import cupic_amd
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(cupic_amd.lib, n)]
    assert not missing, missing


def test_no_cpu_fallback_in_package():
    """The product path must not reach the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "cupic_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.replace("oracle.", "").lower() or f == "build.py", f


def test_params_derivation_matches_oracle():
    import cupic_amd as cp
    from oracle import oracle as O
    for kw in [dict(), dict(ncx=1024, ncy=512, ppc=200, mi=1836.0, kti=0.1, phi_p=-10.0),
               dict(ncx=77, ncy=41, ppc=33, mi=400.0, kti=0.5, phi_p=1.5, ds=0.3, dt=0.02)]:
        p = cp.params(**kw)
        o = O.make_params(**{k: v for k, v in kw.items()})
        assert (p.Lx, p.Ly, p.per_row, p.dtin_e, p.dtin_i, p.eps0) == \
            (o.Lx, o.Ly, o.per_row, o.dtin_e, o.dtin_i, o.eps0)


def test_reference_input_file_reader(tmp_path):
    import cupic_amd as cp
    # the reference's own input/input_data (values copied as data)
    txt = ("#execution configuration\nn_ini = 0;\nn_prev = 0;\nn_save = 100;\n"
           "n_fin = 200000;\nne = 1.0e14;\nTe = 1.0e3;\nbeta = 1.0e-1;\ngamma = 1.0e3;\n"
           "phi_p = -5.0e0;\nncx = 127;\nncy = 1021;\nds = 1.0e-1;\ndt = 1.0e-3;\n")
    f = tmp_path / "input_data"
    f.write_text(txt)
    p, meta = cp.params_from_input_file(str(f))
    assert meta == dict(n_ini=0, n_prev=0, n_save=100, n_fin=200000)
    assert (p.ncx, p.ncy, p.ds, p.dt, p.mi, p.kti) == (127, 1021, 0.1, 1e-3, 1000.0, 0.1)
    # closed forms of src/init.cu:523-556, 612-630, 718-733 (eps0 reduces to n)
    Dl = np.sqrt(8.854e-12 * 1.381e-23 * 1e3 / (1e14 * 1.602e-19**2))
    assert p.n == pytest.approx(1e14 * Dl**3, rel=1e-12)
    assert p.eps0 == pytest.approx(p.n, rel=1e-9)
    assert p.phi_p == pytest.approx(-5.0 * 1.602e-19 / (1.381e-23 * 1e3), rel=1e-12)
    assert p.per_row * p.ncy == 134772
    assert p.jacobi_mode == cp.JACOBI_REFERENCE and p.jacobi_min_sweeps == 1023
    bad = tmp_path / "short"
    bad.write_text("#x\nn_ini = 0;\n")
    with pytest.raises(cp.CupicError) as e:
        cp.params_from_input_file(str(bad))
    assert e.value.code == -7
    with pytest.raises(cp.CupicError):
        cp.params_from_input_file(str(tmp_path / "missing"))


def test_invalid_params_rejected():
    import cupic_amd as cp
    with pytest.raises(cp.CupicError) as e:
        cp.params(ncx=2)
    assert e.value.code == -1
    with pytest.raises(cp.CupicError):
        cp.params(dt=-1.0)
    with pytest.raises(cp.CupicError):
        cp.params(jacobi_mode=7)
    with pytest.raises(cp.CupicError):
        cp.params(direct_method=3)
    for m in (cp.DIRECT_AUTO, cp.DIRECT_DENSE, cp.DIRECT_FFT):
        assert cp.params(direct_method=m).direct_method == m


def test_direct_solve_footprint_by_method():
    """The direct solve's tables: the dense products hold two ncx x ncx and
    one m x m FP64 bases, the FFT path a few rows of twiddles, chirps and
    filter spectra, so on a large grid the dense footprint is far above the
    FFT one, whatever the lengths."""
    import cupic_amd as cp
    kw = dict(ncx=1000, ncy=500, ppc=1, jacobi_mode=cp.POISSON_DIRECT)
    dense = cp.footprint(cp.params(**kw, direct_method=cp.DIRECT_DENSE))
    auto = cp.footprint(cp.params(**kw))            # not powers of two: dense
    fft = cp.footprint(cp.params(**kw, direct_method=cp.DIRECT_FFT))
    assert auto == dense
    assert dense - fft > 8 * (2 * 1000 * 1000 + 499 * 499) * 0.9


def test_footprint_scales_with_particles():
    import cupic_amd as cp
    small = cp.footprint(cp.params())
    big = cp.footprint(cp.params(ncx=1024, ncy=512, ppc=200))
    assert 5e6 < small < 1e8
    # 16 B x 2 buffers x 2 species per particle dominates
    assert big > 64 * 104857600 * 0.95


def test_input_file_n_ini_rewrite(tmp_path):
    """The end of src/main.cu (98-113): line 2 becomes "n_ini = n_fin;"; the
    header and every other line stay as they were."""
    import cupic_amd as cp
    txt = ("#execution configuration\nn_ini = 0;\nn_prev = 0;\nn_save = 100;\n"
           "n_fin = 200000;\nne = 1.0e14;\nTe = 1.0e3;\nbeta = 1.0e-1;\ngamma = 1.0e3;\n"
           "phi_p = -5.0e0;\nncx = 127;\nncy = 1021;\nds = 1.0e-1;\ndt = 1.0e-3;\n")
    f = tmp_path / "input_data"
    f.write_text(txt)
    cp.input_file_set_n_ini(str(f), 200000)
    lines = f.read_text().split("\n")
    want = txt.split("\n")
    assert lines[1] == "n_ini = 200000;"
    assert lines[:1] + lines[2:] == want[:1] + want[2:]
    _, meta = cp.params_from_input_file(str(f))
    assert meta["n_ini"] == 200000
    assert not (tmp_path / "input_data.new").exists()
    with pytest.raises(cp.CupicError) as e:
        cp.input_file_set_n_ini(str(tmp_path / "missing"), 1)
    assert e.value.code == -7


def _run_py(code, tmp_path):
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    return r.returncode, r.stdout + r.stderr


def test_second_library_copy_refused(tmp_path):
    """A process that already maps a copy of libcupic.so (another file, for
    example a variant build) cannot load the package's library beside it,
    and a process that loaded the package refuses to create runs once a
    copy appears.  Runs on the CPU: nothing reaches the GPU."""
    import shutil
    copy = tmp_path / "libcupic_variant.so"
    shutil.copy(os.path.join(ROOT, "cupic_amd", "libcupic.so"), copy)
    code = (f"import ctypes; ctypes.CDLL({str(copy)!r})\n"
            "try:\n"
            "    import cupic_amd\n"
            "except ImportError as e:\n"
            "    print('REFUSED', type(e).__name__, e)\n"
            "else:\n"
            "    print('LOADED')\n")
    rc, out = _run_py(code, tmp_path)
    assert rc == 0 and "REFUSED LibraryCopyError" in out and str(copy) in out, out
    code = ("import cupic_amd as cp\n"
            "assert cp.other_library_copies() == []\n"
            f"import ctypes; ctypes.CDLL({str(copy)!r})\n"
            f"assert cp.other_library_copies() == [{str(copy)!r}], cp.other_library_copies()\n"
            "try:\n"
            "    cp.Run(cp.params())\n"
            "except cp.LibraryCopyError as e:\n"
            "    print('REFUSED', e)\n")
    rc, out = _run_py(code, tmp_path)
    assert rc == 0 and "REFUSED" in out, out
    # the oracle's library is test infrastructure, not a copy
    code = ("from oracle import oracle as O; O.lib()\n"
            "import cupic_amd as cp\n"
            "print('COPIES', cp.other_library_copies())\n")
    rc, out = _run_py(code, tmp_path)
    assert rc == 0 and "COPIES []" in out, out


def test_capacity_bound_and_large_runs():
    """Slot indices and offsets are int32: a species buffer holds at most
    2e9 slots, and runs well past 2^29 slots (where a 4 * slot int32 word
    index would overflow) are accepted with their full footprint."""
    import cupic_amd as cp
    # 1024 x 2048 cells, 300 ppc: 629 M particles per species, cap ~ 708 M slots
    p = cp.params(ncx=1024, ncy=2048, ppc=300)
    n = p.per_row * p.ncy
    assert n > 2 ** 29
    b = cp.footprint(p)
    assert b > 64 * n
    with pytest.raises(cp.CupicError) as e:
        cp.footprint(p, capacity=2_000_000_001)
    assert e.value.code == -1 and "capacity" in str(e.value)
    with pytest.raises(cp.CupicError):
        cp.footprint(p, capacity=n - 1)       # below the initial load
    assert cp.footprint(p, capacity=2_000_000_000) > 64 * 2_000_000_000
    # the kernels address slots through float4 pointers, never as 4 * slot words
    src = open(os.path.join(ROOT, "cupic_amd", "csrc", "k_cc.hip")).read()
    assert "[4 * pos]" not in src and "[4 * (b + u + k)]" not in src
