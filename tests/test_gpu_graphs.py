"""Imaginary-time relaxation replayed as HIP graphs against the same relaxation launched step
by step: the kernels and their arguments are the same, so energies, step counts and states must
agree exactly.  Every replay is followed by the host's own D2H copy and stopping-rule check, so
replays alternate with eager work as they do in use."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

import libgrid_amd as LG
from conftest import rel_l2
from libgrid_amd import workloads as W

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def relax(lg, shape, step, psi, v, prec, graphs, tol=1e-10, max_steps=20000, g=0.05):
    nx, ny, nz = shape
    w = lg.grid_wf_alloc(nx, ny, nz, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"g")
    LG.set_cgrid(w.contents.grid, psi)
    lg.cgrid_set_precision(w.contents.grid, prec)
    r = lg.rgrid_alloc(nx, ny, nz, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(r, v)
    lg.grid_hip_graphs(1 if graphs else 0)
    e = C.c_double()
    lg.grid_hip_sync()
    t0 = time.perf_counter()
    steps = lg.grid_wf_gpe_relax(w, r, g, 0.1, tol, max_steps, C.byref(e))
    lg.grid_hip_sync()
    sec = time.perf_counter() - t0
    out = LG.get_cgrid(w.contents.grid)
    lg.grid_wf_free(w)
    lg.rgrid_free(r)
    lg.grid_hip_graphs(1)
    return e.value, steps, out, sec


@pytest.mark.parametrize("shape,prec", [((1, 128, 128), b"\x01"), ((32, 32, 32), b"\x00"), ((64, 64, 64), b"\x01")])
def test_graph_replay_equals_eager(lg, shape, prec):
    case(f"graph replay vs eager {shape} prec {prec[0]}")
    rng = np.random.default_rng(3)
    step = 0.3
    x = [W.axes(n, step) for n in shape]
    r2 = x[0][:, None, None] ** 2 + x[1][None, :, None] ** 2 + x[2][None, None, :] ** 2
    v = 0.02 * r2 - 0.3 * np.exp(-r2 / 4.0)
    psi = (np.exp(-r2 / 8.0) * (1 + 0.1 * rng.normal(size=shape))).astype(np.complex128)
    ge, gs, gpsi, gsec = relax(lg, shape, step, psi, v, prec, True)
    ee, es, epsi, esec = relax(lg, shape, step, psi, v, prec, False)
    print(f"graphs: {gs} steps {gsec:.3f} s E {ge:.15f}; eager: {es} steps {esec:.3f} s E {ee:.15f}")
    assert gs == es
    assert ge == ee
    np.testing.assert_array_equal(gpsi, epsi)


def relax_extrapolated(lg, shape, step, psi, v, prec, graphs, stride=64, tol=1e-10, max_steps=40000):
    nx, ny, nz = shape
    w = lg.grid_wf_alloc(nx, ny, nz, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"x")
    LG.set_cgrid(w.contents.grid, psi)
    lg.cgrid_set_precision(w.contents.grid, prec)
    r = lg.rgrid_alloc(nx, ny, nz, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(r, v)
    lg.grid_hip_graphs(1 if graphs else 0)
    e = C.c_double()
    nx_ = C.c_long()
    steps = lg.grid_wf_gpe_relax_extrapolated(w, r, 0.0, 0.1, tol, max_steps, C.byref(e), stride, C.byref(nx_))
    out = LG.get_cgrid(w.contents.grid)
    lg.grid_wf_free(w)
    lg.rgrid_free(r)
    lg.grid_hip_graphs(1)
    return e.value, steps, nx_.value, out


@pytest.mark.parametrize("shape,prec", [((1, 128, 128), b"\x01"), ((1, 192, 192), b"\x01"), ((32, 64, 32), b"\x00")])
def test_graph_replay_equals_eager_extrapolated(lg, shape, prec):
    """The extrapolated relaxation (snapshots, reductions and extrapolations between replayed
    batches; the straggler phase of the batched solver runs it on 2-D FP32 grids) gives the same
    steps, extrapolations, energy and state with graphs as without."""
    case(f"extrapolated relaxation graph replay vs eager {shape} prec {prec[0]}")
    # 3-D: a 16-bohr y axis keeps the wells 6 bohr apart both ways round the periodic box (a
    # numpy relaxation of this case extrapolates first after 320 steps; at 32^3 and 0.3 bohr
    # the wells couple across the boundary and the tail is too fast for the rule to fire)
    step = 0.2 if shape[0] == 1 else 0.25
    x = [W.axes(n, step) for n in shape]
    r2 = x[0][:, None, None] ** 2 + x[1][None, :, None] ** 2 + x[2][None, None, :] ** 2
    # two nearly degenerate wells: a slow imaginary-time tail that the extrapolation shortens
    xx = x[1][None, :, None] + 0 * x[2][None, None, :] + 0 * x[0][:, None, None]
    v = -0.3 * np.exp(-((xx - 3.0) ** 2 + (r2 - xx**2)) / 2.0) - 0.2999 * np.exp(-((xx + 3.0) ** 2 + (r2 - xx**2)) / 2.0)
    psi = np.exp(-r2 / 20.0).astype(np.complex128)
    ge, gs, gx, gpsi = relax_extrapolated(lg, shape, step, psi, v, prec, True)
    ee, es, ex, epsi = relax_extrapolated(lg, shape, step, psi, v, prec, False)
    print(f"graphs: {gs} steps, {gx} extrapolations, E {ge:.15f}; eager: {es} steps, {ex} extrapolations, E {ee:.15f}")
    assert gx > 0
    assert (gs, gx) == (es, ex)
    assert ge == ee
    np.testing.assert_array_equal(gpsi, epsi)


@pytest.mark.parametrize("shape", [(16, 32, 24), (8, 12, 64)])
def test_merged_relaxation_matches_oracle_steps(lg, orc, shape):
    """A linear 3-D relaxation (g = 0) runs its batches with merged kinetic half-steps.  With
    tol = 0 it never stops early, so after 40 steps (batches of 16, 16 and 8) the state must be
    the oracle's 40 whole reference steps, each renormalised: K(t/2) V K(t/2) composed of
    grid_wf_propagate and grid_wf_normalize (src/grid_wf.c:379-432, :694-700)."""
    case(f"merged relaxation vs oracle, {shape}, 40 steps")
    step, dt = 0.35, 0.1
    x = [W.axes(n, step) for n in shape]
    r2 = x[0][:, None, None] ** 2 + x[1][None, :, None] ** 2 + x[2][None, None, :] ** 2
    v = 0.02 * r2 - 0.3 * np.exp(-r2 / 4.0)
    rng = np.random.default_rng(41)
    psi = (np.exp(-r2 / 6.0) * (1 + 0.2 * rng.normal(size=shape)) * np.exp(0.3j * x[2][None, None, :])).astype(np.complex128)
    nx, ny, nz = shape
    w = lg.grid_wf_alloc(nx, ny, nz, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"m")
    LG.set_cgrid(w.contents.grid, psi)
    r = lg.rgrid_alloc(nx, ny, nz, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(r, v)
    e = C.c_double()
    steps = lg.grid_wf_gpe_relax(w, r, 0.0, dt, 0.0, 40, C.byref(e))
    got = LG.get_cgrid(w.contents.grid)
    lg.grid_wf_free(w)
    lg.rgrid_free(r)
    a = np.ascontiguousarray(psi.copy())
    norm0 = np.sum(np.abs(a) ** 2) * step**3
    orc.gpe_step_inplace(a, v, 0.0, step, 1.0, (0, 0, 0), 0.0, normalize=True, target=1.0)  # normalise only
    for _ in range(40):
        orc.gpe_step_inplace(a, v, 0.0, step, 1.0, (0, 0, 0), -1j * dt, normalize=True, target=1.0)
    err = rel_l2(got, a)
    eo = orc.gpe_energy(a, v, 0.0, step, 1.0)
    print(f"{shape}: {steps} steps, rel L2 vs oracle {err:.2e}; E {e.value:.15f} oracle {eo:.15f} (start norm {norm0:.3f})")
    assert steps == 40
    assert err < 1e-12
    assert abs(e.value - eo) < 1e-12 * abs(eo)


@pytest.mark.parametrize("shape,prec", [((32, 64, 32), b"\x00"), ((64, 64, 64), b"\x01")])
def test_merged_relaxation_equals_whole_steps(lg, shape, prec):
    """The same relaxation to 1e-10 (FP64) or 1e-8 (FP32) with merged kinetic half-steps and with
    whole steps: the same stopping step and energy, and states equal to rounding."""
    case(f"merged vs whole-step relaxation {shape} prec {prec[0]}")
    step = 0.25
    x = [W.axes(n, step) for n in shape]
    r2 = x[0][:, None, None] ** 2 + x[1][None, :, None] ** 2 + x[2][None, None, :] ** 2
    v = 0.02 * r2 - 0.3 * np.exp(-r2 / 4.0)
    psi = np.exp(-r2 / 8.0).astype(np.complex128)
    tol = 1e-10 if prec == b"\x00" else 1e-8
    out = {}
    for merged in (1, 0):
        lg.grid_hip_merged_relaxation(merged)
        try:
            ge, gs, gpsi, gsec = relax(lg, shape, step, psi, v, prec, True, tol=tol, g=0.0)
        finally:
            lg.grid_hip_merged_relaxation(1)
        out[merged] = (ge, gs, gpsi, gsec)
    (e1, s1, p1, t1), (e0, s0, p0, t0) = out[1], out[0]
    err = rel_l2(p1, p0)
    print(f"merged: {s1} steps {t1:.3f} s E {e1:.15f}; whole steps: {s0} steps {t0:.3f} s E {e0:.15f}; rel L2 {err:.2e}")
    if prec == b"\x00":
        assert s1 == s0
        assert abs(e1 - e0) < 1e-13
        assert err < 1e-10
    else:
        assert abs(e1 - e0) < 3e-8
        assert err < 1e-4
