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
  // FP32 sin / cos of an FP64 argument: error in FP32 ulp of the result (ulp of 2^-24 below
  // |result| = 2^-24 scale kept at the FP32 ulp of 1e-3 so that zeros do not dominate), and the
  // absolute error, against FP32 sin / cos of x reduced mod 2 pi in FP64 and rounded to FP32
  double fs = 0, fc = 0, fa = 0, old_a = 0;
  for (int i = 0; i < 2000000; ++i) {
    double x = i < 1000 ? (i - 500) * 0.7853981633974483 * 0.5 : std::uniform_real_distribution<double>(-2000.0, 2000.0)(g);
    float s, c;
    lg::sincos_f32(x, &s, &c);
    long double rs = sinl((long double)x), rc = cosl((long double)x);
    auto fulp = [](long double r) {
      float a = std::fmax((float)std::fabs((double)r), 1e-3f);
      return (double)(std::nextafter(a, INFINITY) - a);
    };
    fs = std::fmax(fs, (double)std::fabs((long double)s - rs) / fulp(rs));
    fc = std::fmax(fc, (double)std::fabs((long double)c - rc) / fulp(rc));
    fa = std::fmax(fa, (double)std::fmax(std::fabs((long double)s - rs), std::fabs((long double)c - rc)));
    const double twopi = 6.283185307179586476925286766559;
    const double red = x - twopi * std::rint(x * (1.0 / twopi));
    float so, co;
    so = std::sin((float)red); co = std::cos((float)red);
    old_a = std::fmax(old_a, (double)std::fmax(std::fabs((long double)so - rs), std::fabs((long double)co - rc)));
  }
  // arguments formed as a product at the call site (the kernels pass c_im * V): the finiteness
  // test must not be folded with the product into an fma
  double pe = 0;
  for (int i = 0; i < 100000; ++i) {
    volatile double ci = std::uniform_real_distribution<double>(-0.1, 0.1)(g);
    volatile double vr = std::uniform_real_distribution<double>(-50.0, 50.0)(g);
    double s, c;
    lg::sincos_f64(ci * vr, &s, &c);
    pe = std::fmax(pe, std::fabs(s * s + c * c - 1.0));
  }
  double sn, cn, sm, cm;
  lg::sincos_f64(INFINITY, &sn, &cn);
  lg::sincos_f64(std::nan(""), &sm, &cm);
  std::printf("%g %d %d\n", pe, (int)(std::isnan(sn) && std::isnan(sm)), (int)(std::isnan(cn) && std::isnan(cm)));
  // exp at and beyond the ends of its range, and in the subnormal results
  double sub = 0;
  for (int i = 0; i < 20000; ++i) {
    double y = std::uniform_real_distribution<double>(-745.0, -708.5)(g);
    // error in units of the smallest subnormal, the spacing of every result here
    sub = std::fmax(sub, (double)(std::fabs((long double)lg::exp_f64(y) - expl((long double)y)) / 4.9406564584124654e-324L));
  }
  float fsn, fcn;
  lg::sincos_f32(std::nan(""), &fsn, &fcn);
  std::printf("%g %d %d %d %d %d %d %d\n", sub, (int)std::isnan(lg::exp_f64(std::nan(""))),
              (int)std::isinf(lg::exp_f64(709.79)), (int)std::isinf(lg::exp_f64(1e6)),
              (int)(lg::exp_f64(-745.2) == 0.0 && lg::exp_f64(-1e6) == 0.0),
              (int)(std::isfinite(lg::exp_f64(709.78)) && lg::exp_f64(709.78) > 1.7e308),
              (int)(lg::exp_f64(-745.1) > 0.0), (int)(std::isnan(fsn) && std::isnan(fcn)));
  float sz, cz, sq, cq;
  lg::sincos_f32(0.0, &sz, &cz);
  lg::sincos_f32(-3.0 * 1.5707963267948966, &sq, &cq);
  std::printf("%g %g %g %g %g %g %g %g\n", fs, fc, fa, old_a, sz, cz, sq, cq);
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
    es, ec, ee, ew, s0, c0, e0, eunder = map(float, out[:8])
    pe, nan_s, nan_c = map(float, out[8:11])
    assert pe < 1e-15 and nan_s == 1 and nan_c == 1, (pe, nan_s, nan_c)
    sub, *flags = map(float, out[11:19])
    assert sub <= 1.0, sub  # subnormal results within one spacing of exp
    assert all(f == 1 for f in flags), flags  # NaN, overflow, underflow, last finite, first subnormal
    fs, fc, fa, old_a, sz, cz, sq, cq = map(float, out[19:])
    # FP32 sin / cos of an FP64 argument (the FP32 potential step's phase)
    assert fs < 2.0 and fc < 2.0, (fs, fc)
    assert fa < 1.2e-7 and fa < old_a, (fa, old_a)
    assert sz == 0.0 and cz == 1.0
    assert abs(sq - 1.0) < 1e-7 and abs(cq) < 1e-7  # -3 pi/2: quadrant 1 of a negative n
    # near zeros of sin / cos the ulp of the result is tiny; a few ulp of absolute error remain
    assert es < 4.0, es
    assert ec < 4.0, ec
    assert ee < 3.0, ee
    assert ew < 1e-14, ew  # |x| up to 1e6: absolute error still ~ 1e-16 * |x| / 2^20 scale
    assert s0 == 0.0 and c0 == 1.0 and e0 == 1.0 and eunder == 0.0
