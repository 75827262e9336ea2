"""The injection clock's closed form (cupic_amd/csrc/cupic_clock.h, used by
the device kernel k_clock) equals the chain of FP64 adds of the reference's
host loop (src/cc.cu:179, 247) bit for bit.  The header is compiled for the
host with the project's compiler into a small checker (tests/clock_check.cpp)
that compares the two on random clocks and intervals, binade crossings, ties,
oversized intervals, subnormal clocks and overflowing segment lists."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    pytest.skip("hipcc not available")


def test_clock_closed_form_equals_fp64_chain(tmp_path):
    exe = tmp_path / "clock_check"
    subprocess.run([hipcc(), "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "cupic_amd", "csrc"),
                    os.path.join(ROOT, "tests", "clock_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    words = r.stdout.split()
    assert words[-1] == "0" and int(words[1]) > 40000 and int(words[3]) > 40_000_000
