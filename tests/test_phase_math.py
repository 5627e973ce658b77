"""Accuracy of the compact FP64 exp / sincos used by the potential-phase kernels
(libgrid_amd/csrc/hip/phase_math.hpp), compiled for the host and compared with x87
extended-precision libm (expl / sinl / cosl) on seeded random arguments."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include "phase_math.hpp"
static double ulp_err(double got, long double ref) {
  double r = (double)ref;
  if (r == 0.0) return got == 0.0 ? 0.0 : 1e30;
  double u = std::nextafter(std::fabs(r), INFINITY) - std::fabs(r);
  return (double)std::fabs((long double)got - ref) / u;
}
int main() {
  std::mt19937_64 g(1234);
  double es = 0, ec = 0, ee = 0, ew = 0;
  for (int i = 0; i < 2000000; ++i) {
    double x = std::uniform_real_distribution<double>(-2000.0, 2000.0)(g);
    double s, c;
    lg::sincos_f64(x, &s, &c);
    es = std::fmax(es, ulp_err(s, sinl((long double)x)));
    ec = std::fmax(ec, ulp_err(c, cosl((long double)x)));
    double y = std::uniform_real_distribution<double>(-700.0, 700.0)(g);
    ee = std::fmax(ee, ulp_err(lg::exp_f64(y), expl((long double)y)));
    double w = std::uniform_real_distribution<double>(-1e6, 1e6)(g);
    lg::sincos_f64(w, &s, &c);
    ew = std::fmax(ew, std::fmax(std::fabs(s - (double)sinl((long double)w)), std::fabs(c - (double)cosl((long double)w))));
  }
  double s0, c0;
  lg::sincos_f64(0.0, &s0, &c0);
  std::printf("%g %g %g %g %g %g %g %g\n", es, ec, ee, ew, s0, c0, lg::exp_f64(0.0), lg::exp_f64(-800.0));
  return 0;
}
"""


def test_compact_exp_sincos_accuracy():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "h.cpp")
        exe = os.path.join(d, "h")
        open(src, "w").write(HARNESS)
        subprocess.run(["g++", "-O2", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=fast", "-I",
                        os.path.join(ROOT, "libgrid_amd", "csrc", "hip"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    es, ec, ee, ew, s0, c0, e0, eunder = map(float, out)
    # near zeros of sin / cos the ulp of the result is tiny; a few ulp of absolute error remain
    assert es < 4.0, es
    assert ec < 4.0, ec
    assert ee < 3.0, ee
    assert ew < 1e-14, ew  # |x| up to 1e6: absolute error still ~ 1e-16 * |x| / 2^20 scale
    assert s0 == 0.0 and c0 == 1.0 and e0 == 1.0 and eunder == 0.0
