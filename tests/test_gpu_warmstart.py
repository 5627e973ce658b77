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
    """Three held-out 256^3 potentials: network start + FP32 relaxation + FP64 polish (plain, with
    the FP64 stage extrapolated, and with both stages extrapolated) against the plain FP64
    relaxation from the Gaussian."""
    case("warm start 256^3 held-out 4..6")
    modes = ("gauss64", "net32", "net32e", "net32ee")
    recs = WS.evaluate(lg, net, 3, WS.FINE, modes, first=4)
    s = WS.summarize(recs, modes)
    print(s)
    assert s["energy_spread_max"] < 2e-10
    for m in modes[1:]:
        assert s[m]["speedup_total"] > 1.5, m


@pytest.mark.timeout(900)
def test_fast_path_speedup_on_heldout_subset(lg, net):
    """The time-to-ground-state target on the fixed held-out subset (every 8th of the 200 held-out
    256^3 potentials, 25 of them): the fast path (network start, FP32 relaxation to 3e-8 Eh, FP64
    polish to 1e-10 Eh, vector extrapolation in both stages) against plain FP64 relaxation from the
    Gaussian start, both timed including building the start, both judged by the converged
    energy."""
    case(f"warm start fast path vs plain FP64, 256^3 held-out subset {WS.HELDOUT_SUBSET}")
    modes = ("gauss64", "net32ee")
    recs = WS.evaluate(lg, net, 0, WS.FINE, modes, indices=WS.HELDOUT_SUBSET)
    s = WS.summarize(recs, modes)
    print(s)
    assert s["gauss64"]["converged"] == s["count"] and s["net32ee"]["converged"] == s["count"]
    # both stop on the rule's estimate at tol = 1e-10 (spread over all 200: 2.2e-10)
    assert s["energy_spread_max"] < 3e-10
    assert s["net32ee"]["speedup_total"] >= 3.0


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


def test_extrapolated_relaxation_64(lg):
    """grid_wf_gpe_relax_extrapolated against the plain relaxation from the Gaussian start on
    three slow held-out 64^3 potentials: the same ground-state energy (both to 1e-10 Eh), fewer
    steps in total, and the same result bit for bit with and without HIP-graph replay."""
    case("extrapolated relaxation 64^3, held-out 0, 2, 15")
    n = 64
    prm = W.wells_3d_params(WS.HELDOUT_SEED, 16, WS.BOX)
    e = C.c_double()
    nx = C.c_long()
    plain_steps = extra_steps = 0
    for i in (0, 2, 15):
        p = WS.Problem(lg, n, W.wells_3d(prm[i], n, WS.BOX / n))
        try:
            p.start(None)
            sp = lg.grid_wf_gpe_relax(p.w, p.r, 0.0, WS.DT, 1e-10, 200000, C.byref(e))
            ep = lg.grid_wf_gpe_energy(p.w, p.r, 0.0)
            runs = []
            for graphs in (1, 0):
                lg.grid_hip_graphs(graphs)
                p.start(None)
                sx = lg.grid_wf_gpe_relax_extrapolated(p.w, p.r, 0.0, WS.DT, 1e-10, 200000, C.byref(e), 64,
                                                       C.byref(nx))
                runs.append((sx, nx.value, LG.get_cgrid(p.w.contents.grid)))
            lg.grid_hip_graphs(1)
            ex = lg.grid_wf_gpe_energy(p.w, p.r, 0.0)
            print(f"potential {i}: plain {sp} steps E {ep:.13f}; extrapolated {runs[0][0]} steps "
                  f"({runs[0][1]} extrapolations) E {ex:.13f}")
            assert sp < 200000 and runs[0][0] < 200000
            assert abs(ex - ep) < 2e-10
            assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
            np.testing.assert_array_equal(runs[0][2], runs[1][2])
            plain_steps += sp
            extra_steps += runs[0][0]
        finally:
            lg.grid_hip_graphs(1)
            p.free()
    assert extra_steps < 0.7 * plain_steps
