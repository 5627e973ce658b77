"""hgrid_sum_norm2_noalloc, the allocation-free variant of hgrid_sum_norm2 (its partials live in a
fixed per-device array instead of caller scratch), against hgrid_sum_norm2 with library-allocated
scratch on the same inputs: bit-exact (the same kernels in the same order), on two calls in a row,
the second with a smaller input, in FP64 and FP32 storage; sizes outside 1 .. 2^34 are refused
with an error before anything is launched."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def device_array(lg, host):
    p = C.c_void_p()
    assert lg.hgrid_malloc(C.byref(p), host.nbytes) == 0
    assert lg.hgrid_memcpy_h2d(p, host.ctypes.data, host.nbytes) == 0
    return p


def fetch(lg, p):
    v = np.zeros(1)
    assert lg.hgrid_memcpy_d2h(v.ctypes.data, p, 8) == 0
    return float(v[0])


@pytest.mark.parametrize("prec,dtype", [(0, np.complex128), (1, np.complex64)])
def test_noalloc_norm_matches_scratch_version(lg, prec, dtype):
    rng = np.random.default_rng(5)
    big = (rng.standard_normal((1 << 24) + 7) + 1j * rng.standard_normal((1 << 24) + 7)).astype(dtype)
    small = (rng.standard_normal(3001) + 1j * rng.standard_normal(3001)).astype(dtype)
    out_a, out_b = C.c_void_p(), C.c_void_p()
    assert lg.hgrid_malloc(C.byref(out_a), 16) == 0 and lg.hgrid_malloc(C.byref(out_b), 16) == 0
    for host in (big, small):
        n = host.size
        case(f"sum_norm2 noalloc vs scratch: n {n}, prec {prec}")
        d = device_array(lg, host)
        scr = C.c_void_p()
        assert lg.hgrid_malloc(C.byref(scr), 8 * lg.hgrid_reduce_scratch(n)) == 0
        assert lg.hgrid_sum_norm2(out_a, d, prec, n, None, scr) == 0
        assert lg.hgrid_sum_norm2_noalloc(out_b, d, prec, n, None) == 0
        assert lg.hgrid_sync() == 0
        a, b = fetch(lg, out_a), fetch(lg, out_b)
        ref = float(np.sum(np.abs(host.astype(np.complex128)) ** 2))
        print(f"n {n}: scratch {a!r} noalloc {b!r} numpy {ref!r}")
        assert a == b
        assert abs(b - ref) <= 1e-12 * ref
        lg.hgrid_free(scr)
        lg.hgrid_free(d)
    lg.hgrid_free(out_a)
    lg.hgrid_free(out_b)


def test_noalloc_norm_refuses_out_of_range_sizes(lg):
    case("sum_norm2 noalloc: sizes outside 1 .. 2^34 are refused")
    out = C.c_void_p()
    assert lg.hgrid_malloc(C.byref(out), 16) == 0
    for n in (0, -5, (1 << 34) + 1):
        assert lg.hgrid_sum_norm2_noalloc(out, out, 0, n, None) == -1
        assert b"hgrid_sum_norm2_noalloc" in lg.hgrid_last_error()
    lg.hgrid_free(out)
