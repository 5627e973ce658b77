"""Batched 2-D imaginary-time ground states in one persistent launch (complex64 storage, FP64
reductions, vector extrapolation inside the launch) against the FP64 path on the same problems:
the FP64 reference relaxes each problem with grid_wf_gpe_relax (g = 0), which is pinned to the
oracle in test_gpu_parity.py, and a few problems are relaxed by the CPU oracle directly."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

import libgrid_amd as LG
from libgrid_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 0.1


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def make(lg, problems):
    wfs, pots = [], []
    for psi, v, prm in problems:
        n = prm["n"]
        w = lg.grid_wf_alloc(1, n, n, prm["step"], 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"b2d")
        LG.set_cgrid(w.contents.grid, psi)
        r = lg.rgrid_alloc(1, n, n, prm["step"], LG.boundary("rgrid_value_outside_periodic"), None, b"v")
        LG.set_rgrid(r, v)
        wfs.append(w)
        pots.append(r)
    return wfs, pots


def run_batch(lg, wfs, pots, tol, max_steps):
    ns = len(wfs)
    warr = (LG.WF * ns)(*wfs)
    parr = (LG.RG * ns)(*pots)
    e = (C.c_double * ns)()
    st = (C.c_long * ns)()
    secs = lg.grid_wf_batch2d_relax(ns, warr, parr, DT, tol, max_steps, e, st)
    return secs, np.array(e[:]), np.array(st[:])


MAX_STEPS = 200000  # nearly degenerate wells relax over tens of thousands of steps at dt = 0.1


def fp64_reference(lg, problem, tol=1e-10, max_steps=MAX_STEPS):
    psi, v, prm = problem
    n = prm["n"]
    w = lg.grid_wf_alloc(1, n, n, prm["step"], 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"ref")
    LG.set_cgrid(w.contents.grid, psi)
    r = lg.rgrid_alloc(1, n, n, prm["step"], LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(r, v)
    e = C.c_double()
    steps = lg.grid_wf_gpe_relax(w, r, 0.0, DT, tol, max_steps, C.byref(e))
    out = LG.get_cgrid(w.contents.grid)
    lg.grid_wf_free(w)
    lg.rgrid_free(r)
    return e.value, steps, out


@pytest.mark.parametrize("stride", [64, 0])
def test_batch2d_small_vs_oracle_and_fp64(lg, orc, stride):
    problems = W.wells_2d(seed=7, count=6, sizes=(64, 96))
    case(f"batch2d small (extrapolation stride {stride}): sizes {[p[2]['n'] for p in problems]}")
    wfs, pots = make(lg, problems)
    lg.grid_batch2d_set_extrapolation(stride)
    try:
        secs, e, st = run_batch(lg, wfs, pots, 1e-8, MAX_STEPS)
    finally:
        lg.grid_batch2d_set_extrapolation(64)
    print(f"extrapolations: {lg.grid_batch2d_last_extrapolations()}")
    if stride == 0:
        assert lg.grid_batch2d_last_extrapolations() == 0
    for i, (psi, v, prm) in enumerate(problems):
        e64, s64, st64 = fp64_reference(lg, problems[i])
        got = LG.get_cgrid(wfs[i].contents.grid)
        overlap = abs(np.vdot(got.ravel(), st64.ravel())) * prm["step"] ** 2
        print(f"solve {i}: n={prm['n']} E32={e[i]:.10f} (step {st[i]}) E64={e64:.10f} ({s64} steps) "
              f"|dE|={abs(e[i] - e64):.2e} |<32|64>|={overlap:.8f}")
        assert abs(e[i] - e64) < 1e-7
        assert overlap > 1 - 1e-6
    # the oracle relaxes the first problem in FP64 on the CPU (grid_wf_propagate + normalize)
    psi, v, prm = problems[0]
    a = np.ascontiguousarray(psi.copy())
    e_prev = None
    for k in range(40000):
        orc.gpe_step_inplace(a, np.ascontiguousarray(v), 0.0, prm["step"], 1.0, (0, 0, 0), -1j * DT)
        eo = orc.gpe_energy(a, np.ascontiguousarray(v), 0.0, prm["step"], 1.0)
        if e_prev is not None and abs(eo - e_prev) < 1e-11:
            break
        e_prev = eo
    print(f"oracle: E = {eo:.12f} after {k + 1} steps; batched FP32 {e[0]:.12f}")
    assert abs(e[0] - eo) < 1e-7
    for w in wfs:
        lg.grid_wf_free(w)
    for r in pots:
        lg.rgrid_free(r)


@pytest.mark.timeout(600)
def test_batch2d_256_solves(lg):
    problems = W.wells_2d(seed=0, count=256)
    sizes = [p[2]["n"] for p in problems]
    case(f"batch2d 256 solves, sizes histogram {np.unique(sizes, return_counts=True)}")
    wfs, pots = make(lg, problems)
    ns = len(wfs)
    report = lg.grid_batch2d_autotune(ns, (LG.WF * ns)(*wfs), (LG.RG * ns)(*pots), DT, 20).decode()
    print(report)
    tol = 1e-8
    t0 = time.perf_counter()
    secs, e, st = run_batch(lg, wfs, pots, tol, MAX_STEPS)
    wall = time.perf_counter() - t0
    conv = int(np.sum(st < MAX_STEPS))
    print(f"batched: {ns} solves in {secs:.3f} s GPU ({wall:.3f} s wall incl. packing) = {ns / secs:.1f} solves/s; "
          f"converged {conv}/{ns}; steps min/median/max {st.min()}/{int(np.median(st))}/{st.max()}")
    p2n, p2s = C.c_long(), C.c_double()
    lg.grid_batch2d_last_phase2(C.byref(p2n), C.byref(p2s))
    print(f"straggler phase: {p2n.value} solves continued on the multi-kernel FP32 path in {p2s.value:.3f} s; "
          f"extrapolations inside the launch: {lg.grid_batch2d_last_extrapolations()}")
    # every solve finishes inside the one persistent launch
    assert p2n.value == 0
    t_s = (C.c_double * ns)()
    t_e = (C.c_double * ns)()
    assert lg.grid_batch2d_solve_times(t_s, t_e) == ns
    dur = np.array(t_e[:]) - np.array(t_s[:])
    span = max(t_e[:])
    print(f"device clock: makespan {span:.3f} s, busy {dur.sum():.3f} WG-s; slowest solves (n, steps, s): "
          f"{[(sizes[i], int(st[i]), round(float(dur[i]), 3)) for i in np.argsort(-dur)[:5]]}")
    errs, s64s = [], []
    t1 = time.perf_counter()
    for i in range(ns):
        e64, s64, _ = fp64_reference(lg, problems[i])
        errs.append(abs(e[i] - e64))
        s64s.append(s64)
    ref_time = time.perf_counter() - t1
    errs = np.array(errs)
    w = int(np.argmax(errs))
    print(f"per-solve |E_fp32 - E_fp64|: max {errs.max():.2e} (solve {w}: n={sizes[w]}, {st[w]} FP32 / {s64s[w]} FP64 "
          f"steps) median {np.median(errs):.2e} mean {errs.mean():.2e}; FP64 one-by-one reference took {ref_time:.1f} s")
    assert conv == ns
    assert max(s64s) < MAX_STEPS
    # FP32 states with FP64 reductions: converged energies agree with FP64 to a few 1e-9
    assert errs.max() < 1e-7
    assert np.median(errs) < 2e-8
    for w in wfs:
        lg.grid_wf_free(w)
    for r in pots:
        lg.rgrid_free(r)
