#!/usr/bin/env python3
"""Benchmark: FFT split-operator steps/s at 256^3 FP64 (one MI355X).

Workload (one "step"): one imaginary-time Gross-Pitaevskii split-operator step on a 256^3
complex128 grid, spacing 0.25 bohr: V_eff = 0.5*0.01*r^2 + 0.05*exp(-r^2/4) + 0.1*|psi|^2
built from the start-of-step state, K(dt/2) V(dt) K(dt/2) with dt = 0.05, renormalisation,
and the energy of the start-of-step state — libgrid's grid_wf_gpe_propagate, i.e. the
reference composition grid_wf_density -> grid_real_to_complex_re -> grid_wf_propagate ->
grid_wf_normalize fused into 9 pencil passes.  Inputs are synthetic (seeded) and resident in
HBM before timing starts.

Prints one JSON line with steps/s, the roofline of the dominant kernel (HIP events on the
library stream) and the CPU oracle's throughput on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "split-op steps/s @256^3 (% HBM roofline); time-to-ground-state at 1e-10 Eh"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def classify_pass(LG, axis, nfft, flags):
    """Name of a pass kind of the fused GPE step from its axis (0 outer, 2 contiguous), number of
    transforms and flags."""
    if axis == 2 and flags & LG.PF_NOSTORE:
        return "z_inv_pre"
    if axis == 2 and nfft == 2 and flags & LG.PF_MID_GPE:
        return "z_gpe"
    if axis == 2 and nfft == 2 and flags & LG.PF_MID_POT:
        return "z_pot"
    if axis == 1 and nfft == 1:
        return "y_fft"
    if axis == 0 and nfft == 2 and flags & LG.PF_MID_KIN:
        return "x_kin"
    return f"axis{axis}_nfft{nfft}_flags{flags:#x}"


def pass_bytes_per_point(name, S):
    """Algorithmic HBM bytes per grid point of one launch of each pass kind of the deferred GPE
    step (S = bytes per stored complex value): the state read and written once, plus what the
    fused pointwise work reads or writes (vext 8 B, V_eff 8 B)."""
    return {"z_inv_pre": S + 8 + 8,  # S read (not stored), vext read, V_eff written
            "y_fft": 2 * S, "x_kin": 2 * S,
            "z_gpe": 2 * S + 8}.get(name)  # V_eff read


def insitu_passes(lib, LG, step_fn, steps, N, S):
    """Time every pass kernel of `steps` real steps with HIP events on the library stream
    (grid_hip_pass_timing) and group the launches by pass kind: average ms per launch, launches
    per step, algorithmic bytes per launch and their rate."""
    cap = 4096
    lib.grid_hip_pass_timing(1)
    for _ in range(steps):
        step_fn()
    arr = {k: (C.c_int * cap)() for k in ("axis", "nfft", "prec")}
    flags, lens, ms = (C.c_uint * cap)(), (C.c_long * cap)(), (C.c_double * cap)()
    k = lib.grid_hip_pass_times(cap, arr["axis"], arr["nfft"], flags, lens, arr["prec"], ms)
    lib.grid_hip_pass_timing(0)
    groups = {}
    for i in range(k):
        groups.setdefault(classify_pass(LG, arr["axis"][i], arr["nfft"][i], flags[i]), []).append(ms[i])
    out = {}
    for name, t in groups.items():
        avg = float(np.mean(t))
        bpp = pass_bytes_per_point(name, S)
        rec = dict(ms=avg, count=len(t) / steps, min_ms=float(np.min(t)), max_ms=float(np.max(t)))
        if bpp is not None:
            rec.update(bytes=bpp * N, gbps=bpp * N / (avg * 1e-3) / 1e9, frac=bpp * N / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS)
        out[name] = rec
    return out


def copy_ceiling(lib, nbytes, reps=10):
    """hgrid_stream_copy (one 16-byte load and store per thread, the fastest copy shape measured on
    MI355X) of one buffer into another: read + write bytes over the average time."""
    a, b = C.c_void_p(), C.c_void_p()
    if lib.hgrid_malloc(C.byref(a), nbytes) or lib.hgrid_malloc(C.byref(b), nbytes):
        raise RuntimeError(lib.hgrid_last_error().decode())
    lib.hgrid_memset(a, 0, nbytes)
    for _ in range(3):
        lib.hgrid_stream_copy(b, a, nbytes)
    import libgrid_amd as LG
    t = LG.HipTimer()
    t.start()
    for _ in range(reps):
        lib.hgrid_stream_copy(b, a, nbytes)
    ms = t.stop() / reps
    t.close()
    lib.hgrid_free(a)
    lib.hgrid_free(b)
    gbps = 2 * nbytes / (ms * 1e-3) / 1e9
    log(f"stream copy {nbytes >> 20} MiB: {ms * 1e3:.1f} us, {gbps:.0f} GB/s")
    return dict(bytes_each_way=nbytes, ms=ms, GBps=gbps, frac=gbps / HBM_PEAK_GBPS)


def fp64_512_bench(lib, LG, steps=10, warmup=3):
    """The FP64 GPE imaginary-time step at 512^3 (spacing 0.125, the same trap and barrier, g 0.1,
    dt 0.05, renormalised): a 2 GiB state, eight times the 256 MB Infinity Cache, so every pass
    streams from HBM.  Steps/s, the step's fraction of the HBM roofline on the bytes it moves, and
    every pass kind timed in place with HIP events."""
    n, step, g, dt = 512, 0.125, 0.1, 0.05
    w = lib.grid_wf_alloc(n, n, n, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"gpe512")
    vr = lib.rgrid_alloc(n, n, n, step, LG.boundary("rgrid_value_outside_periodic"), None, b"vext512")
    x = (np.arange(n) - n // 2) * step
    r2 = (x * x)[:, None, None] + (x * x)[None, :, None] + (x * x)[None, None, :]
    LG.set_rgrid(vr, 0.5 * 0.01 * r2 + 0.05 * np.exp(-r2 / 4.0))
    del r2
    lib.grid_wf_gaussian_start(w, 10.0)
    tstep = LG.c_complex.of(-1j * dt)

    def one():
        lib.grid_wf_gpe_propagate(w, vr, g, tstep, b"\x01")

    for _ in range(warmup):
        one()
    lib.grid_hip_sync()
    t = LG.HipTimer()
    t.start()
    for _ in range(steps):
        one()
    ms = t.stop() / steps
    t.close()
    N = n ** 3
    passes = insitu_passes(lib, LG, one, 3, N, 16)
    actual = sum(v["bytes"] * v["count"] for v in passes.values() if "bytes" in v)
    lib.grid_wf_free(w)
    lib.rgrid_free(vr)
    out = dict(workload="gpe_barrier 512^3 complex128 imaginary-time step (spacing 0.125, dt 0.05, g 0.1, renormalised)",
               ms_per_step=ms, steps_per_s=1000.0 / ms, actual_bytes_per_step=actual,
               actual_GBps=actual / (ms * 1e-3) / 1e9, frac=actual / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, passes=passes)
    log(f"fp64 512^3: {ms:.3f} ms/step, {actual / (ms * 1e-3) / 1e9:.0f} GB/s on {actual / N:.0f} B/pt "
        f"({out['frac']:.3f} of 8 TB/s); passes " +
        ", ".join(f"{k} {v['ms'] * 1e3:.0f} us {v.get('frac', 0):.2f}" for k, v in passes.items()))
    return out


# template instantiation of each pass of the 256^3 FP64 imaginary-time step in the rocprofv3
# summaries (length, contiguous, kernel kind, storage type, wide tiles): the kinds of
# libgrid_amd/csrc/hip/fft_pass.hpp are 8 plain transform (y), 6 real kinetic factors (x),
# 5 imaginary-time GPE (z), 4 start-of-step (z)
PASS_SYMBOL = {name: f"256, {layout}, {kind}, HIP_vector_type<double, 2u>, false>"
               for name, layout, kind in (("y_fft", "false", 8), ("x_kin", "false", 6), ("z_gpe", "true", 5),
                                          ("z_inv_pre", "true", 4))}


def profiled_traffic(pass_name, n):
    """HBM bytes per launch of a pass from the committed PMC summary (FETCH_SIZE + WRITE_SIZE,
    calibrated against the streaming copy), or None when it is not available for this pass."""
    path = os.path.join(ROOT, "profiles", "gpe256_fft_pass_hbm_traffic.txt")
    key = PASS_SYMBOL.get(pass_name)
    if n != 256 or key is None or not os.path.exists(path):
        return None
    for line in open(path):
        if "fft_pass_kernel<HIP_vector_type<double, 2u>, " + key in line:
            parts = line.split()
            return (float(parts[0]) + float(parts[1])) * 1024 * 1024
    return None


def ground_state(lib, LG, psi, vext, prm, n, max_steps):
    """Wall time to |E_n - E_{n-1}| < 1e-10 Eh from the workload's initial state:
    (a) FP64 throughout; (b) FP32 storage until the energy change falls below 1e-7 Eh (the
    FP32 noise floor is ~1e-8), then FP64 to 1e-10; (c) as (b) with vector extrapolation every
    64 steps in the FP64 stage; (d) as (b) with it in the FP32 stage; (e) with it in both.
    Energies of all runs are reported."""
    step, g, dt, tol = prm["step"], prm["g"], prm["dt"], prm["tol"]
    out = {}
    for mode in ("fp64", "fp32_then_fp64", "fp32_then_fp64e", "fp32e_then_fp64", "fp32e_then_fp64e"):
        w = lib.grid_wf_alloc(n, n, n, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"gs")
        vr = lib.rgrid_alloc(n, n, n, step, LG.boundary("rgrid_value_outside_periodic"), None, b"vext")
        LG.set_cgrid(w.contents.grid, psi)
        LG.set_rgrid(vr, vext)
        e = C.c_double()
        # upload and table set-up outside the timed region
        lib.grid_wf_gpe_energy(w, vr, g)
        if mode != "fp64":
            lib.cgrid_set_precision(w.contents.grid, b"\x01")
        lib.grid_hip_sync()
        t0 = time.perf_counter()
        if mode == "fp64":
            steps = lib.grid_wf_gpe_relax(w, vr, g, dt, tol, max_steps, C.byref(e))
            s32 = 0
        else:
            if mode in ("fp32e_then_fp64", "fp32e_then_fp64e"):
                s32 = lib.grid_wf_gpe_relax_extrapolated(w, vr, g, dt, 1e-7, max_steps, C.byref(e), 64, None)
            else:
                s32 = lib.grid_wf_gpe_relax(w, vr, g, dt, 1e-7, max_steps, C.byref(e))
            lib.cgrid_set_precision(w.contents.grid, b"\x00")
            if mode in ("fp32_then_fp64e", "fp32e_then_fp64e"):
                steps = lib.grid_wf_gpe_relax_extrapolated(w, vr, g, dt, tol, max_steps, C.byref(e), 64, None)
            else:
                steps = lib.grid_wf_gpe_relax(w, vr, g, dt, tol, max_steps, C.byref(e))
        lib.grid_hip_sync()
        el = time.perf_counter() - t0
        efinal = lib.grid_wf_gpe_energy(w, vr, g)
        out[mode] = dict(seconds=el, fp64_steps=steps, fp32_steps=s32, energy=efinal, converged=steps < max_steps)
        log(f"ground state [{mode}]: {el:.3f} s, fp32 steps {s32}, fp64 steps {steps}, E = {efinal:.12f}")
        lib.grid_wf_free(w)
        lib.rgrid_free(vr)
    out["speedup_fp32_then_fp64"] = out["fp64"]["seconds"] / out["fp32_then_fp64"]["seconds"]
    out["speedup_fp32_then_fp64e"] = out["fp64"]["seconds"] / out["fp32_then_fp64e"]["seconds"]
    out["speedup_fp32e_then_fp64"] = out["fp64"]["seconds"] / out["fp32e_then_fp64"]["seconds"]
    out["speedup_fp32e_then_fp64e"] = out["fp64"]["seconds"] / out["fp32e_then_fp64e"]["seconds"]
    out["tol_Eh"] = tol
    return out


def warm_start_bench(lib, count, subset=False):
    """Wall time to a 1e-10 Eh ground state on held-out 256^3 wells potentials: plain Gaussian
    start with FP64 throughout, against the trained network's start with FP32 relaxation
    (to 3e-8) then FP64 (to 1e-10): with vector extrapolation in both stages (net32ee, the
    fast path), in the FP64 stage only (net32e) and in neither (net32), and against the network
    start in FP64 alone (net64).  All timings include building the start state."""
    from libgrid_amd import warmstart as WS

    if not os.path.exists(WS.WEIGHTS):
        log(f"warm start: no trained network at {WS.WEIGHTS}; skipped")
        return None
    net = lib.grid_surrogate_load(WS.WEIGHTS.encode(), 1)
    modes = ("gauss64", "net32ee", "net32e", "net32", "net64") if not subset else ("gauss64", "net32ee")
    recs = WS.evaluate(lib, net, count, WS.FINE, modes, indices=WS.HELDOUT_SUBSET if subset else None)
    lib.grid_surrogate_free(net)
    summ = WS.summarize(recs, modes)
    log(f"warm start: {json.dumps(summ)}")
    return dict(summary=summ, potentials=[r["index"] for r in recs], seed=WS.HELDOUT_SEED,
                grid=[WS.FINE] * 3, box_bohr=WS.BOX, dt=WS.DT, tol_Eh=1e-10)


def fp32_512_bench(lib, LG, steps=20, warmup=3):
    """The 512^3 complex64 real-time GPE step of the vortex-ring workload (dt 0.02, g 0.1,
    harmonic trap omega 0.1, spacing 0.15): FP64 twiddles, arguments and reductions.  The step's
    cost does not depend on the state, so a Gaussian condensate built on the host stands in for
    the ring (its accuracy run is tests/test_gpu_configs.py).  Steps/s, the step's fraction of the
    HBM roofline on the bytes it moves, and every pass kind timed in place with HIP events."""
    n, step, g, dt, omega = 512, 0.15, 0.1, 0.02, 0.1
    w = lib.grid_wf_alloc(n, n, n, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"ring")
    vr = lib.rgrid_alloc(n, n, n, step, LG.boundary("rgrid_value_outside_periodic"), None, b"trap")
    x = (np.arange(n) - n // 2) * step
    r2 = (x * x)[:, None, None] + (x * x)[None, :, None] + (x * x)[None, None, :]
    LG.set_rgrid(vr, 0.5 * omega * omega * r2)
    del r2
    lib.cgrid_set_precision(w.contents.grid, b"\x01")
    lib.grid_wf_gaussian_start(w, 10.0)

    def one():
        lib.grid_wf_gpe_propagate(w, vr, g, LG.c_complex.of(dt), b"\x00")

    for _ in range(warmup):
        one()
    lib.grid_hip_sync()
    t = LG.HipTimer()
    t.start()
    for _ in range(steps):
        one()
    ms = t.stop() / steps
    t.close()
    N = n ** 3
    passes = insitu_passes(lib, LG, one, 3, N, 8)
    actual = sum(v["bytes"] * v["count"] for v in passes.values() if "bytes" in v)
    lib.grid_wf_free(w)
    lib.rgrid_free(vr)
    out = dict(workload="vortex_ring 512^3 complex64 real-time GPE step (dt 0.02, g 0.1)", ms_per_step=ms,
               steps_per_s=1000.0 / ms, actual_bytes_per_step=actual, actual_GBps=actual / (ms * 1e-3) / 1e9,
               frac=actual / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, passes=passes)
    log(f"fp32 512^3: {ms:.3f} ms/step, {1000 / ms:.1f} steps/s, {out['actual_GBps']:.0f} GB/s on {actual / N:.0f} B/pt "
        f"({out['frac']:.3f} of 8 TB/s); passes " +
        ", ".join(f"{k} {v['ms'] * 1e3:.0f} us {v.get('frac', 0):.2f}" for k, v in passes.items()))
    return out


def propagators_bench(lib, LG, steps=20, warmup=3):
    """grid_wf_propagate at 256^3 FP64 (spacing 0.25, real time dt 0.05, complex cgrid
    potential of the barrier-trap workload) for each FFT propagator: the plain 2nd-order step,
    its k-space cutoff (CFFT, kmax at half the Nyquist |k|), the 4th-order step and the
    absorbing boundary (layers of 32 points on every face).  Model bytes of the 2nd-order step:
    400 B/pt."""
    n, step, dt = 256, 0.25, 0.05
    x = (np.arange(n) - n // 2) * step
    r2 = (x * x)[:, None, None] + (x * x)[None, :, None] + (x * x)[None, None, :]
    pot = (0.5 * 0.01 * r2 + 0.05 * np.exp(-r2 / 4.0)).astype(np.complex128)
    del r2
    vp = lib.cgrid_alloc(n, n, n, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
    LG.set_cgrid(vp, pot)
    del pot
    kmax = 0.5 * np.sqrt(3.0) * np.pi / step
    cases = (("fft2", LG.WF_2ND_ORDER_FFT, None), ("cfft2", LG.WF_2ND_ORDER_CFFT, None),
             ("fft4", LG.WF_4TH_ORDER_FFT, None), ("fft2_absorbing", LG.WF_2ND_ORDER_FFT, (32, 223) * 3))
    out = {}
    for name, prop, bounds in cases:
        w = lib.grid_wf_alloc(n, n, n, step, 1.0, LG.WF_PERIODIC_BOUNDARY, prop, b"wf")
        w.contents.kmax = kmax
        lib.grid_wf_gaussian_start(w, 4.0)
        if bounds:
            lib.grid_wf_boundary(w, None, *bounds)
        for _ in range(warmup):
            lib.grid_wf_propagate(w, vp, LG.c_complex.of(dt))
        lib.grid_hip_sync()
        t = LG.HipTimer()
        t.start()
        for _ in range(steps):
            lib.grid_wf_propagate(w, vp, LG.c_complex.of(dt))
        ms = t.stop() / steps
        t.close()
        lib.grid_wf_free(w)
        out[name] = dict(ms_per_step=ms, steps_per_s=1000.0 / ms,
                         model_frac=400 * n ** 3 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS)
        log(f"propagator {name:15s} {ms:.3f} ms/step ({1000 / ms:.0f} steps/s)")
    lib.cgrid_free(vp)
    out["workload"] = "grid_wf_propagate 256^3 complex128, real time dt 0.05, complex cgrid potential"
    return out


def batch2d_bench(lib, LG, count=256, tol=1e-8, dt=0.1):
    """The batched 2-D workload: 256 seeded problems (n in 64..256, 3-8 Gaussian wells, spacing
    0.2) relaxed to tol in one persistent launch (autotuned tile)."""
    from libgrid_amd import workloads as W

    problems = W.wells_2d(seed=0, count=count)
    wfs, pots = [], []
    for psi, v, prm in problems:
        nn = prm["n"]
        w = lib.grid_wf_alloc(1, nn, nn, prm["step"], 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"b2d")
        LG.set_cgrid(w.contents.grid, psi)
        r = lib.rgrid_alloc(1, nn, nn, prm["step"], LG.boundary("rgrid_value_outside_periodic"), None, b"v")
        LG.set_rgrid(r, v)
        wfs.append(w)
        pots.append(r)
    warr, parr = (LG.WF * count)(*wfs), (LG.RG * count)(*pots)
    report = lib.grid_batch2d_autotune(count, warr, parr, dt, 20).decode()
    e = (C.c_double * count)()
    st = (C.c_long * count)()
    t0 = time.perf_counter()
    secs = lib.grid_wf_batch2d_relax(count, warr, parr, dt, tol, 200000, e, st)
    wall = time.perf_counter() - t0
    p2n, p2s = C.c_long(), C.c_double()
    lib.grid_batch2d_last_phase2(C.byref(p2n), C.byref(p2s))
    ts, te = (C.c_double * count)(), (C.c_double * count)()
    lib.grid_batch2d_solve_times(ts, te)
    busy = float(sum(te[i] - ts[i] for i in range(count)))
    cus = lib.hgrid_b2d_cus()
    steps = np.array(st[:])
    for w in wfs:
        lib.grid_wf_free(w)
    for r in pots:
        lib.rgrid_free(r)
    out = dict(workload=f"{count} x 2-D wells (n in 64..256, spacing 0.2, dt {dt}, tol {tol}), one persistent launch",
               solves_per_s=count / secs, launch_s=secs, wall_s_incl_packing=wall, steps_median=int(np.median(steps)),
               steps_max=int(steps.max()), extrapolations=int(lib.grid_batch2d_last_extrapolations()),
               straggler_solves=int(p2n.value), straggler_s=p2s.value, busy_workgroup_s=busy,
               utilisation=busy / (cus * secs) if cus > 0 else None, autotune=report)
    log(f"batch2d: {count / secs:.1f} solves/s ({secs:.3f} s, one launch, {out['extrapolations']} extrapolations, "
        f"{p2n.value} stragglers), steps median {out['steps_median']} max {out['steps_max']}, busy {busy:.2f} WG-s, "
        f"utilisation {out['utilisation']:.3f}")
    return out


def cpu_baseline(psi, vext, prm, steps):
    from oracle import oracle as orc

    nthreads = os.cpu_count() or 1
    orc.set_num_threads(nthreads)
    a = np.ascontiguousarray(psi.copy())
    step, g, dt = prm["step"], prm["g"], prm["dt"]
    orc.gpe_step_inplace(a, vext, g, step, 1.0, (0, 0, 0), -1j * dt)  # warm-up (first touch, twiddles)
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.gpe_step_inplace(a, vext, g, step, 1.0, (0, 0, 0), -1j * dt)
        orc.gpe_energy(a, vext, g, step, 1.0)
    el = time.perf_counter() - t0
    return dict(value=steps / el, unit="steps/s", cores=orc.num_threads(), kind="port",
                sample=f"{steps} GPE imaginary-time steps (+energy) of the same 256^3 FP64 workload after 1 warm-up "
                       f"step, oracle/oracle.c (C/OpenMP restatement, own mixed-radix FFT; FFTW absent), "
                       f"{el:.1f} s")


def summarize_rocpd(stats_dir, fetch_dir, write_dir, command):
    """Text summary of rocprofv3 runs of this benchmark: kernel time statistics, and HBM bytes per
    dispatch from FETCH_SIZE / WRITE_SIZE scaled by what the same counters read for the streaming
    copy kernel (which moves a known number of bytes: one grid read and one grid written)."""
    import glob
    import sqlite3

    def db(d):
        f = sorted(glob.glob(os.path.join(d, "**", "*.db"), recursive=True))
        if not f:
            raise FileNotFoundError(f"no rocpd database under {d}")
        return sqlite3.connect(f[0])

    out = [f"command: {command}", ""]
    if stats_dir:
        c = db(stats_dir)
        out.append("kernel statistics (rocprofv3 --kernel-trace --stats): calls, total us, average us, % of GPU time")
        for name, calls, tot, avg, pct in c.execute("select name, total_calls, total_duration, average, percentage from top_kernels"):
            out.append(f"  {calls:5d} {tot:12.1f} {avg:10.2f} {pct:6.2f}%  {name}")
        out.append("")
    if fetch_dir and write_dir:
        vals = {}
        for tag, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
            c = db(d)
            for name, v, n in c.execute("select kernel_name, avg(value), count(*) from counters_collection where counter_name = ? "
                                        "group by kernel_name", (tag,)):
                vals.setdefault(name, {})[tag] = (v, n)
        # the copy runs on one 256 MiB grid and on a 2 GiB buffer: calibrate on the 256 MiB
        # dispatches (the smaller of the two groups of counter values, eight times apart)
        grid_kib = 256.0 ** 3 * 16 / 1024.0
        factors = {}
        for tag, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
            v = sorted(r[0] for r in db(d).execute("select value from counters_collection where counter_name = ? and "
                                                   "kernel_name like '%k_stream_copy%'", (tag,)))
            if v:
                small = [x for x in v if x < 2.83 * v[0]]
                factors[tag] = grid_kib / (sum(small) / len(small))
        if len(factors) == 2:
            ff, fw = factors["FETCH_SIZE"], factors["WRITE_SIZE"]
            out.append(f"calibration: stream copy of one 256^3 complex128 grid ({grid_kib:.0f} KiB each way) reads "
                       f"FETCH_SIZE {grid_kib / ff:.0f} KiB, WRITE_SIZE {grid_kib / fw:.0f} KiB -> factors fetch x{ff:.3f}, "
                       f"write x{fw:.3f}")
            out.append("HBM bytes per dispatch (calibrated): MiB read, MiB written, dispatches  kernel")
            for name in sorted(vals):
                if "FETCH_SIZE" not in vals[name] or "WRITE_SIZE" not in vals[name]:
                    continue
                r = vals[name]["FETCH_SIZE"][0] * ff / 1024.0
                w = vals[name]["WRITE_SIZE"][0] * fw / 1024.0
                out.append(f"  {r:10.1f} {w:10.1f} {vals[name]['FETCH_SIZE'][1]:5d}  {name}")
    return "\n".join(out)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-rocpd":
        ap = argparse.ArgumentParser()
        ap.add_argument("--summarize-rocpd", action="store_true")
        ap.add_argument("--stats", default=None)
        ap.add_argument("--fetch", default=None)
        ap.add_argument("--write", default=None)
        ap.add_argument("--command", default="")
        a = ap.parse_args()
        print(summarize_rocpd(a.stats, a.fetch, a.write, a.command))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gs", action="store_true", help="skip the time-to-ground-state runs")
    ap.add_argument("--gs-max-steps", type=int, default=20000)
    ap.add_argument("--no-rows", action="store_true", help="skip the FP32 512^3, batched 2-D and propagator-variant measurements")
    ap.add_argument("--rows", default="fp64_512,fp32_512,batch2d,propagators",
                    help="comma-separated subset of the other rows to measure")
    ap.add_argument("--ws-count", type=int, default=-1,
                    help="held-out 256^3 wells potentials for the warm-start time-to-ground-state comparison: "
                         "the first N in all five modes, or -1 (default) for the fixed subset of every 8th "
                         "of the 200 (25 potentials) in the plain and fast modes, 0 to skip")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch relaxation batches eagerly instead of replaying HIP graphs (for profiling: "
                         "rocprofv3 --kernel-trace has crashed inside hipGraphLaunch once a process had destroyed "
                         "one graph and instantiated another)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library drives one GPU per process and has no sharded "
              f"path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import libgrid_amd as LG
    from libgrid_amd import workloads as W

    lib = LG.load()
    lib.grid_hip_enable(b"\x01", 1, None)
    if args.no_graphs:
        lib.grid_hip_graphs(0)
    log(f"device {lib.hgrid_device_name().decode()}")
    n = args.n
    t0 = time.perf_counter()
    psi, vext, prm = W.gpe_barrier(seed=0, n=n)
    log(f"inputs {prm} built in {time.perf_counter() - t0:.1f} s")
    step, g, dt = prm["step"], prm["g"], prm["dt"]
    w = lib.grid_wf_alloc(n, n, n, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"psi")
    vr = lib.rgrid_alloc(n, n, n, step, LG.boundary("rgrid_value_outside_periodic"), None, b"vext")
    LG.set_cgrid(w.contents.grid, psi)
    LG.set_rgrid(vr, vext)
    tstep = LG.c_complex.of(-1j * dt)
    for _ in range(args.warmup):
        lib.grid_wf_gpe_propagate(w, vr, g, tstep, b"\x01")
    lib.grid_hip_sync()
    timer = LG.HipTimer()
    wall0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        lib.grid_wf_gpe_propagate(w, vr, g, tstep, b"\x01")
    ms_total = timer.stop()
    lib.grid_hip_sync()
    wall = time.perf_counter() - wall0
    ms_step = ms_total / args.steps
    energy = lib.grid_wf_gpe_last_energy(w)
    log(f"{args.steps} steps: {ms_step:.3f} ms/step (events), wall {wall / args.steps * 1e3:.3f} ms/step, E = {energy:.12f}")

    # every pass kernel of further real steps, timed in place with HIP events on the library stream
    N = n**3
    kinds = insitu_passes(lib, LG, lambda: lib.grid_wf_gpe_propagate(w, vr, g, tstep, b"\x01"), args.steps, N, 16)
    dname, d = max(((k, v) for k, v in kinds.items() if "bytes" in v), key=lambda kv: kv[1]["ms"] * kv[1]["count"])
    sum_ms = sum(v["ms"] * v["count"] for v in kinds.values())
    actual_bytes = sum(v["bytes"] * v["count"] for v in kinds.values() if "bytes" in v)
    model_bytes = 400 * N  # (4 FFTs x 3 sweeps x 2 x 16 B + 16 B potential) per point
    for k, v in kinds.items():
        log(f"pass {k:10s} {v['ms'] * 1e3:8.1f} us x{v['count']:.0f}/step" +
            (f"  {v['gbps']:6.0f} GB/s ({v['frac']:.3f})" if "gbps" in v else ""))
    copies = {"256MiB": copy_ceiling(lib, N * 16), "2GiB": copy_ceiling(lib, 8 * N * 16)}
    traffic = profiled_traffic(dname, n)

    gs = None
    if not args.no_gs:
        gs = ground_state(lib, LG, psi, vext, prm, n, args.gs_max_steps)

    ws = None
    if not args.no_gs and args.ws_count != 0:
        ws = warm_start_bench(lib, args.ws_count, subset=args.ws_count < 0)

    rows = {}
    if not args.no_rows:
        which = set(args.rows.split(","))
        if "fp32_512" in which:
            rows["fp32_512"] = fp32_512_bench(lib, LG)
        if "batch2d" in which:
            rows["batch2d"] = batch2d_bench(lib, LG)
        if "propagators" in which:
            rows["propagators"] = propagators_bench(lib, LG)
        if "fp64_512" in which:
            rows["fp64_512"] = fp64_512_bench(lib, LG)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(psi, vext, prm, args.cpu_steps)
        log(f"cpu baseline {cpu['value']:.4f} steps/s on {cpu['cores']} threads")

    result = {
        "metric": METRIC,
        "value": 1000.0 / ms_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded GPE barrier-trap state, random smoothed phase)",
        "config": {"workload": f"gpe_barrier {n}^3 complex128 imaginary-time split-operator step (dt 0.05, g 0.1, "
                               f"renormalised, start-of-step energy)", "grid": [n, n, n], "step_bohr": step,
                   "passes_per_step": 8},
        "roofline": {"bound": "hbm", "kernel": f"fft_pass {dname}", "achieved": d["gbps"], "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": d["gbps"] / HBM_PEAK_GBPS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": d["bytes"], "avg_launch_ms": d["ms"]},
        # the step on the bytes it moves (264 B/pt: the deferred z transform cancels two z sweeps
        # per step); the 400 B/pt model of a 12-sweep step is kept for reference only
        "step_roofline": {"actual_bytes_per_step": actual_bytes,
                          "achieved_GBps": actual_bytes / (ms_step * 1e-3) / 1e9,
                          "frac": actual_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                          "model_bytes_per_step_12_sweeps": model_bytes,
                          "sum_of_pass_ms": sum_ms,
                          "note": "the 256 MiB state fits the 256 MB Infinity Cache, so part of the traffic is served "
                                  "from it; other_rows.fp64_512 is the same step at 512^3 (2 GiB), streaming from HBM"},
        "passes": kinds,
        "copy_ceiling": copies,
        "energy": energy,
        "time_to_ground_state": gs,
        "warm_start": ws,
        "other_rows": rows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    lib.grid_wf_free(w)
    lib.rgrid_free(vr)


if __name__ == "__main__":
    main()
