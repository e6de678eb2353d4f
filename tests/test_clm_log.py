"""The library's logarithm (clm_log in csrc/clm_physics.h) and the powers built
on it, against the C library in extended precision.

clm_log is built for the host from the same header (exact division and the
C library's frexp in place of the device's reciprocal and frexp builtins)
and compared with logl on 8 million arguments: uniform in log-space over
[e^-40, e^40], clustered at 1, and across the reduction boundary √½ and the
binade [0.5, 2).  Bar: 1 ulp (the atanh-series construction gives about
0.85).  Special values: log 0 = -inf, log inf = inf, log of a negative or
NaN = NaN, log 1 = +0, denormals.  clm_pow(x, y) = exp(y log x) is held to
1.5 |y log x| + 2 ulp (log's ulp and the product's rounding, scaled by the
exponent, plus exp's own).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "climaland.jl_amd", "csrc")

HARNESS = r"""
#include "clm_physics.h"
#include <cmath>
#include <cstdio>
#include <random>
static double ulps(double got, long double ref) {
    double r = (double)ref;
    double u = std::nextafter(std::fabs(r), INFINITY) - std::fabs(r);
    return (double)(((long double)got - ref) / u);
}
int main() {
    std::mt19937_64 g(7);
    auto uni = [&]() { return (double)(g() >> 11) * 0x1.0p-53; };
    double mx = 0.0, xm = 0.0;
    for (int i = 0; i < 8000000; ++i) {
        double x;
        switch (i % 4) {
            case 0: x = std::exp(-40.0 + 80.0 * uni()); break;
            case 1: x = 1.0 + (uni() - 0.5) * 1e-3; break;
            case 2: x = 0.5 + 1.5 * uni(); break;
            default: x = 0.70710678118654752 * (1.0 + (uni() - 0.5) * 1e-6); break;
        }
        double e = std::fabs(ulps(clm::clm_log(x), logl((long double)x)));
        if (e > mx) { mx = e; xm = x; }
    }
    std::printf("log_max_ulp %.6f %.17g\n", mx, xm);
    double pm = 0.0;
    for (int i = 0; i < 1000000; ++i) {
        double x = std::exp(-20.0 + 25.0 * uni()), y = -3.0 + 6.0 * uni();
        long double ref = powl((long double)x, (long double)y);
        double bound = 1.5 * std::fabs(y * std::log(x)) + 2.0;
        double e = std::fabs(ulps(clm::clm_pow(x, y), ref)) / bound;
        if (e > pm) pm = e;
    }
    std::printf("pow_max_ratio %.6f\n", pm);
    std::printf("special %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", clm::clm_log(0.0),
                clm::clm_log(INFINITY), clm::clm_log(-1.0), clm::clm_log(NAN), clm::clm_log(1.0),
                clm::clm_log(4.9406564584124654e-324), clm::clm_log(2.2250738585072014e-308));
    std::printf("signbit_log1 %d\n", (int)std::signbit(clm::clm_log(1.0)));
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("clmlog")
    src, exe = d / "h.cpp", d / "h"
    src.write_text(HARNESS)
    r = subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-I", CSRC,
                        "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {ln.split()[0]: ln.split()[1:] for ln in out.splitlines()}


def test_log_within_one_ulp(harness):
    assert float(harness["log_max_ulp"][0]) <= 1.0, harness["log_max_ulp"]


def test_pow_error_bound(harness):
    assert float(harness["pow_max_ratio"][0]) <= 1.0


def test_log_special_values(harness):
    s = [float(v) for v in harness["special"]]
    assert s[0] == -np.inf and s[1] == np.inf
    assert np.isnan(s[2]) and np.isnan(s[3])
    assert s[4] == 0.0 and harness["signbit_log1"] == ["0"]
    assert s[5] == pytest.approx(np.log(5e-324), rel=1e-15)
    assert s[6] == pytest.approx(np.log(2.2250738585072014e-308), rel=1e-15)
