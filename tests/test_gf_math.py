"""The library's own FP64 log (goofit_amd/csrc/gf_math.h), the -log of the likelihood
metrics in every sweep in place of the device library's: built for the host with the same
correctly rounded operations (tests/cpp/math_check.cpp) and compared with long double logl
over every binary exponent, around 1, where e ln 2 and log m cancel, and on subnormal
inputs.

Bar: below one unit in the last place everywhere (a wrong series coefficient, reduction
constant or reconstruction step shows up as errors of many units), the IEEE results at the
special values, and the nearest double at every power of two."""
import math
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("gfmath") / "math_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "goofit_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "math_check.cpp"), "-o", exe], check=True)
    return subprocess.run([exe, "2000000", "11"], check=True, capture_output=True, text=True).stdout


@pytest.mark.parametrize("name", ["log_range", "log_near_one", "log_cancelling", "log_subnormal_inputs"])
def test_below_one_ulp(report, name):
    m = re.search(name + r" n=(\d+) max_ulp=([\d.]+) worst_x=(\S+) not_nearest=(\d+)", report)
    n, max_ulp, not_nearest = int(m.group(1)), float(m.group(2)), int(m.group(4))
    assert n > 1_000_000
    assert max_ulp < 1.0, m.group(0)
    assert not_nearest < 0.1 * n, m.group(0)  # correctly rounded in the large majority of cases


def test_special_values(report):
    sp = {}
    for line in report.splitlines():
        m = re.match(r"special x=(\S+) log=(\S+)", line)
        if m:
            sp[m.group(1)] = float(m.group(2))
    assert sp["0"] == -math.inf and sp["-0"] == -math.inf
    assert sp["1"] == 0.0
    assert sp["inf"] == math.inf
    assert math.isnan(sp["-1"]) and math.isnan(sp["-inf"]) and math.isnan(sp["nan"])
    assert sp["4.9406564584124654e-324"] == math.log(5e-324)
    assert sp["2"] == math.log(2.0) and sp["0.5"] == math.log(0.5)
    assert sp["1.7976931348623157e+308"] == math.log(1.7976931348623157e308)


def test_log_of_powers_of_two(report):
    m = re.search(r"log_powers_of_two not_nearest=(\d+)", report)
    assert m and int(m.group(1)) == 0


def test_nll_as_one_logarithm(report):
    """gfmath::nll_add / nll_total, the per-lane NLL of every likelihood kernel: the sum of
    -log(v) over 2e6 values across every binary exponent (subnormals included, with zeros,
    negatives and NaNs that add 0), over 4 likelihood-like values and over 1000 values,
    against the compensated long double sum of -logl(v): within 1e-16 relative, a small
    fraction of one rounding of the summed magnitudes (a dropped exponent, a wrong ln 2 part
    or a factor not renormalised is off by many orders)."""
    rows = re.findall(r"nll_product trial=(\d) n=(\d+) rel_err=(\S+) ulps_of_scale=(\S+)", report)
    assert len(rows) == 3
    for trial, n, rel, ulps in rows:
        assert float(rel) < 1e-16, (trial, n, rel)
        assert float(ulps) < 1.0, (trial, n, ulps)
    m = re.search(r"nll_special empty=(\S+) with_inf=(\S+)", report)
    assert float(m.group(1)) == 0.0 and float(m.group(2)) == -math.inf
    m = re.search(r"nll_special exact_pair=(\S+)", report)
    assert float(m.group(1)) == 0.0  # 0.125 * 8 = 1 exactly


def test_nll_integer_exponent_form(report):
    """The overload that keeps the exponent in an int (the matrix-core kernels') tracks
    the double form exactly: the same mantissa bits and exponent value after each of 1e6
    factors across every binary exponent, zeros and NaNs included."""
    m = re.search(r"nll_int_exponent n=(\d+) mismatches=(\d+)", report)
    assert m and int(m.group(1)) == 1_000_000 and int(m.group(2)) == 0
