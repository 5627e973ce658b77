"""The trained warm-start network (libgrid_amd/data/unet64.bin) against the plain Gaussian
start: both relax the same held-out potentials to 1e-10 Eh with the library's solver, which
decides the energy.  The network start must reach the same ground state in less time."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import libgrid_amd as LG
from libgrid_amd import warmstart as WS
from libgrid_amd import workloads as W

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def net(lg):
    assert os.path.exists(WS.WEIGHTS), "trained network missing: python -m libgrid_amd.warmstart train"
    h = lg.grid_surrogate_load(WS.WEIGHTS.encode(), 1)
    assert h
    yield h
    lg.grid_surrogate_free(h)


def test_network_start_64(lg, net):
    """16 held-out 64^3 potentials, FP64 relaxation from each start."""
    case("warm start 64^3 x 16 held-out")
    modes = ("gauss64", "net64")
    recs = WS.evaluate(lg, net, 16, 64, modes, every=4)
    s = WS.summarize(recs, modes)
    print(s)
    assert s["gauss64"]["converged"] == 16 and s["net64"]["converged"] == 16
    assert s["energy_spread_max"] < 2e-10
    assert s["net64"]["fp64_steps_mean"] < 0.8 * s["gauss64"]["fp64_steps_mean"]
    assert s["net64"]["speedup_total"] > 1.2


def test_network_fp32_start_256(lg, net):
    """Three held-out 256^3 potentials: network start + FP32 relaxation + FP64 polish against
    the plain FP64 relaxation from the Gaussian."""
    case("warm start 256^3 held-out 4..6")
    modes = ("gauss64", "net32")
    recs = WS.evaluate(lg, net, 3, WS.FINE, modes, first=4)
    s = WS.summarize(recs, modes)
    print(s)
    assert s["energy_spread_max"] < 2e-10
    assert s["net32"]["speedup_total"] > 1.5


def test_no_network_falls_back_to_gaussian(lg):
    case("warm start fallback")
    n = 32
    prm = W.wells_3d_params(WS.HELDOUT_SEED, 1, WS.BOX)[0]
    v = W.wells_3d(prm, n, WS.BOX / n)
    p = WS.Problem(lg, n, v)
    try:
        lg.grid_wf_warm_start(p.w, p.r, None, WS.GAUSS_WIDTH)
        a = LG.get_cgrid(p.w.contents.grid)
        lg.grid_wf_gaussian_start(p.w, WS.GAUSS_WIDTH)
        b = LG.get_cgrid(p.w.contents.grid)
        np.testing.assert_array_equal(a, b)
        assert abs(np.sum(np.abs(b) ** 2) * (WS.BOX / n) ** 3 - 1.0) < 1e-12
    finally:
        p.free()
