"""The f64 log of the dB image output (halide-sar-app_amd/csrc/log_f64.h), built for the host from
the same header the device code includes, against the C library's long double logl: within one
unit in the last place over (0, 1] (the range of |z| / max), every binade, near 1 and subnormal
inputs, exact at powers of two and 1, and IEEE special values (img_output.cpp:10-39 takes the log
of |z| / max)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("log") / "log_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-Wall",
                    "-I", os.path.join(ROOT, "halide-sar-app_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "img_math", "log_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    print(out)
    return out


def _sets(out):
    rows = {}
    for m in re.finditer(r"^(\w+) n=(\d+) max_ulp=([\d.]+) max_ulp_vs_libm=([\d.]+)", out, re.M):
        rows[m.group(1)] = (int(m.group(2)), float(m.group(3)), float(m.group(4)))
    return rows


def test_log_within_one_ulp(report):
    rows = _sets(report)
    assert set(rows) == {"unit_interval", "all_binades", "near_one", "subnormal", "powers_of_two"}
    for name, (n, ulp, ulp_libm) in rows.items():
        assert n > 1000, name
        assert ulp < 1.0, (name, ulp)        # faithful: within one ulp of the exact log
        assert ulp_libm <= 1.0, (name, ulp_libm)
    assert rows["powers_of_two"][1] <= 0.5    # k ln 2: correctly rounded


def test_log_special_values(report):
    line = [ln for ln in report.splitlines() if ln.startswith("special")][0]
    assert "log(1)=0 " in line
    assert "log(0)=-inf" in line and "log(-0)=-inf" in line
    assert re.search(r"log\(-1\)=-?nan", line) and re.search(r"log\(nan\)=-?nan", line)
    assert "log(inf)=inf" in line
