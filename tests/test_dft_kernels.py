"""The in-register DFT butterflies of libgrid_amd/csrc/hip/cx.hpp (the building block of every
GPU transform), compiled for the host with hipcc and compared with a brute-force DFT in long
double for every supported length, in FP64 and FP32."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include "cx.hpp"
template <int N, typename C> double check() {
  std::mt19937_64 g(N);
  std::normal_distribution<double> d;
  C x[N];
  long double xr[N], xi[N];
  for (int i = 0; i < N; ++i) { x[i] = lg::mk<C>(d(g), d(g)); xr[i] = x[i].x; xi[i] = x[i].y; }
  lg::dft<N>(x);
  long double num = 0, den = 0;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < N; ++k) {
    long double sr = 0, si = 0;
    for (int n = 0; n < N; ++n) {
      long double a = -2 * pi * (long double)((long)n * k % N) / N;
      sr += xr[n] * cosl(a) - xi[n] * sinl(a);
      si += xr[n] * sinl(a) + xi[n] * cosl(a);
    }
    num += (x[k].x - sr) * (x[k].x - sr) + (x[k].y - si) * (x[k].y - si);
    den += sr * sr + si * si;
  }
  return (double)sqrtl(num / den);
}
#define RUN(N) std::printf("%d %g %g\n", N, check<N, double2>(), check<N, float2>());
int main() { RUN(2) RUN(3) RUN(4) RUN(6) RUN(8) RUN(12) RUN(16) RUN(24) RUN(32) RUN(48) RUN(64) RUN(96) RUN(192) return 0; }
"""


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc absent")
def test_register_dft_matches_brute_force():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "h.hip")
        exe = os.path.join(d, "h")
        open(src, "w").write(HARNESS)
        subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "libgrid_amd", "csrc", "hip"),
                        src, "-o", exe], check=True, capture_output=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [list(map(float, l.split())) for l in out if l.strip()]
    assert len(rows) == 13
    for n, e64, e32 in rows:
        assert e64 < 4e-16, (n, e64)
        assert e32 < 2.5e-7, (n, e32)
