"""The benchmark's roofline line reads its HBM traffic from the committed rocprofv3 summary: every
pass of the 256^3 FP64 step must be found there under the kernel instantiation that the library
launches for it, and its traffic must be the pass's algorithmic bytes (no re-reads)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

N = 256


@pytest.mark.parametrize("name", ["y_fft", "x_kin", "z_gpe", "z_inv_pre"])
def test_profiled_traffic_found_and_algorithmic(name):
    traffic = bench.profiled_traffic(name, N)
    assert traffic is not None, f"{name}: no line for {bench.PASS_SYMBOL[name]} in the traffic summary"
    algorithmic = bench.pass_bytes_per_point(name, 16) * N ** 3
    assert abs(traffic / algorithmic - 1.0) < 0.01


def test_kernel_statistics_name_every_pass():
    path = os.path.join(ROOT, "profiles", "gpe256_kernel_stats.txt")
    text = open(path).read()
    for name, sym in bench.PASS_SYMBOL.items():
        assert "fft_pass_kernel<HIP_vector_type<double, 2u>, " + sym in text, name
