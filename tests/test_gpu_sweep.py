"""Parity sweeps over every supported axis length on each axis.  The fused GPE step in the three
storage precisions, for imaginary and real time: three steps of grid_wf_gpe_propagate (with
renormalisation) against the oracle's composition of reference calls from the same FP64 input.
The propagators with a complex potential (2nd order in the three precisions, CFFT and 4th
order in FP64), the same potential grid serving every precision in turn.
It reaches every pass kernel kind and geometry the library picks from the shape: one-thread,
two-stage and three-stage pencils, wide outer-axis tiles, partial workgroups, radix-3 lengths and
squeezed 2-D shapes.

Bars: FP64 relative L2 <= 1e-12 (measured <= 1.1e-15); FP32 butterflies <= 2e-6 (measured
<= 4.1e-7); FP32 storage with FP64 arithmetic <= 1e-6 (measured <= 1.6e-7)."""
import sys

import numpy as np
import pytest

import libgrid_amd as LG
from conftest import rel_l2

pytestmark = pytest.mark.gpu

LENGTHS = [2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024]
SHAPES = [s for L in LENGTHS for s in ((L, 4, 6), (6, L, 4), (4, 6, L))] + [
    (1, 64, 8), (64, 1, 8), (48, 96, 16), (512, 8, 8), (1024, 4, 4), (8, 1024, 4), (4, 4, 1024), (1, 512, 12),
    (12, 512, 1)]
BARS = {b"\x00": 1e-12, b"\x01": 2e-6, b"\x02": 1e-6}


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_gpe_steps_every_length(lg, orc, shape):
    print(f"[case] sweep {shape}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(sum(shape))
    step = 0.3
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1 + 1.0
    psi = psi / np.sqrt(np.sum(np.abs(psi) ** 2) * step ** 3)
    vext = rng.uniform(0, 1, shape)
    for time in (-0.05j, 0.03):
        ref = psi.copy()
        for _ in range(3):
            orc.gpe_step_inplace(ref, vext, 0.3, step, 1.0, (0, 0, 0), time, normalize=True)
        for prec, bar in BARS.items():
            w = lg.grid_wf_alloc(*shape, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"w")
            LG.set_cgrid(w.contents.grid, psi)
            if prec != b"\x00":
                lg.cgrid_set_precision(w.contents.grid, prec)
            vr = lg.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
            LG.set_rgrid(vr, vext)
            for _ in range(3):
                lg.grid_wf_gpe_propagate(w, vr, 0.3, LG.c_complex.of(time), 1)
            err = rel_l2(LG.get_cgrid(w.contents.grid), ref)
            lg.grid_wf_free(w)
            lg.rgrid_free(vr)
            assert err < bar, f"{shape} precision {prec[0]} t={time}: rel L2 {err:.3e}"


def kmax_for(shape, step, frac):
    """frac times the largest |k| on the grid (Nyquist along every non-trivial axis)."""
    return frac * np.sqrt(sum((np.pi / step) ** 2 for n in shape if n > 1))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_propagators_every_length(lg, orc, shape):
    """grid_wf_propagate with a complex potential over the same shapes: the 2nd-order FFT step in
    the three precisions (complex-potential and complex kinetic kernels), its k-space cutoff
    (CFFT, kmax = 0.6 of the corner |k|) and the 4th-order step, both in FP64."""
    print(f"[case] propagators {shape}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(7 + sum(shape))
    step, mass, time = 0.3, 1.2, 0.02
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1 + 1.0
    pot = (rng.uniform(0, 1, shape) + 0.05j * rng.uniform(0, 1, shape)).astype(np.complex128)
    kmax = kmax_for(shape, step, 0.6)
    cases = [(LG.WF_2ND_ORDER_FFT, p, bar, None) for p, bar in BARS.items()]
    cases += [(LG.WF_2ND_ORDER_CFFT, b"\x00", 1e-12, kmax), (LG.WF_4TH_ORDER_FFT, b"\x00", 1e-12, None)]
    vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
    LG.set_cgrid(vp, pot)
    for prop, prec, bar, km in cases:
        ref = psi.copy()
        for _ in range(3):
            if prop == LG.WF_2ND_ORDER_FFT:
                orc.wf_propagate_inplace(ref, pot, step, mass, (0, 0, 0), time)
            elif prop == LG.WF_2ND_ORDER_CFFT:
                orc.wf_propagate_cfft_inplace(ref, pot, step, mass, (0, 0, 0), time, km)
            else:
                orc.wf_propagate_4th_inplace(ref, pot, step, mass, (0, 0, 0), time)
        w = lg.grid_wf_alloc(*shape, step, mass, LG.WF_PERIODIC_BOUNDARY, prop, b"w")
        if km is not None:
            w.contents.kmax = km
        LG.set_cgrid(w.contents.grid, psi)
        if prec != b"\x00":
            lg.cgrid_set_precision(w.contents.grid, prec)
        for _ in range(3):
            lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        err = rel_l2(LG.get_cgrid(w.contents.grid), ref)
        lg.grid_wf_free(w)
        assert err < bar, f"{shape} propagator {prop} precision {prec[0]}: rel L2 {err:.3e}"
    lg.cgrid_free(vp)


FFT_BARS = {b"\x00": 1e-12, b"\x01": 1e-6, b"\x02": 2e-7}


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fft_every_length(lg, orc, shape):
    """cgrid_fft and cgrid_inverse_fft_norm over the same shapes in the three storage precisions,
    against the oracle of the input the grid holds (rounded to complex64 for the FP32 storage
    modes)."""
    print(f"[case] fft {shape}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(3 + sum(shape))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    step = 0.3
    for prec, bar in FFT_BARS.items():
        x = a if prec == b"\x00" else a.astype(np.complex64).astype(np.complex128)
        g = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"g")
        LG.set_cgrid(g, a)
        if prec != b"\x00":
            lg.cgrid_set_precision(g, prec)
        lg.cgrid_fft(g)
        fx = orc.fft3d(x, -1)
        err_f = rel_l2(LG.get_cgrid(g), fx)
        lg.cgrid_inverse_fft_norm(g)
        err_i = rel_l2(LG.get_cgrid(g), x)
        lg.cgrid_free(g)
        assert err_f < bar, f"{shape} precision {prec[0]} forward: rel L2 {err_f:.3e}"
        assert err_i < 2 * bar, f"{shape} precision {prec[0]} round trip: rel L2 {err_i:.3e}"


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_energies_every_length(lg, orc, shape):
    """The reductions over the same shapes (FP64): norm, potential expectation, grid_wf_energy by
    both methods (FFT kinetic energy, and the finite-difference default), the GP energy, and
    the energy the fused GPE step records for its start state."""
    print(f"[case] energies {shape}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(5 + sum(shape))
    step, mass, g = 0.3, 1.0, 0.3
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape) + 2.0
    v = rng.uniform(0, 1, shape)
    w = lg.grid_wf_alloc(*shape, step, mass, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"w")
    LG.set_cgrid(w.contents.grid, a)
    vr = lg.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(vr, v)
    checks = [
        ("norm", lg.grid_wf_norm(w), orc.integral_of_square(a, step)),
        ("expectation", lg.grid_grid_expectation_value(w.contents.grid, vr), orc.grid_expectation(a, v, step)),
    ]
    lg.grid_wf_analyze_method(b"\x01")
    checks.append(("energy fft", lg.grid_wf_energy(w, vr), orc.energy_fft(a, v, step, mass)))
    lg.grid_wf_analyze_method(b"\x00")
    checks.append(("energy fd", lg.grid_wf_energy(w, vr),
                   orc.kinetic_energy_fd_periodic(a, step, mass) + orc.grid_expectation(a, v, step)))
    checks.append(("gpe energy", lg.grid_wf_gpe_energy(w, vr, g), orc.gpe_energy(a, v, g, step, mass)))
    lg.grid_wf_gpe_propagate(w, vr, g, LG.c_complex.of(-0.01j), 1)
    checks.append(("step energy", lg.grid_wf_gpe_last_energy(w), orc.gpe_energy(a, v, g, step, mass)))
    lg.grid_wf_free(w)
    lg.rgrid_free(vr)
    for name, got, ref in checks:
        assert abs(got - ref) <= 1e-12 * abs(ref), f"{shape} {name}: {got!r} against {ref!r}"


@pytest.mark.parametrize("shape", [s for s in SHAPES if min(s) > 1], ids=lambda s: "x".join(map(str, s)))
def test_merged_relaxation_every_length(lg, orc, shape):
    """Linear 3-D relaxation (g = 0) over the 3-D shapes: its batches merge the kinetic half-steps
    of consecutive steps (kernels of the merged stages, full-step kinetic tables).  With tol = 0
    it runs all 20 steps (a batch of 16 and one of 4); the state must be the oracle's 20 whole
    reference steps, each renormalised, and the energy the oracle's."""
    import ctypes as C

    print(f"[case] merged relaxation {shape}", flush=True)
    sys.stdout.flush()
    step, dt = 0.35, 0.1
    x = [(np.arange(n) - n // 2) * step for n in shape]
    r2 = x[0][:, None, None] ** 2 + x[1][None, :, None] ** 2 + x[2][None, None, :] ** 2
    v = 0.02 * r2 - 0.3 * np.exp(-r2 / 4.0)
    rng = np.random.default_rng(41 + sum(shape))
    psi = (np.exp(-r2 / 6.0) * (1 + 0.2 * rng.normal(size=shape)) + 0.05).astype(np.complex128)
    w = lg.grid_wf_alloc(*shape, step, 1.0, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"m")
    LG.set_cgrid(w.contents.grid, psi)
    r = lg.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(r, v)
    e = C.c_double()
    steps = lg.grid_wf_gpe_relax(w, r, 0.0, dt, 0.0, 20, C.byref(e))
    got = LG.get_cgrid(w.contents.grid)
    lg.grid_wf_free(w)
    lg.rgrid_free(r)
    a = np.ascontiguousarray(psi.copy())
    orc.gpe_step_inplace(a, v, 0.0, step, 1.0, (0, 0, 0), 0.0, normalize=True, target=1.0)  # normalise only
    for _ in range(20):
        orc.gpe_step_inplace(a, v, 0.0, step, 1.0, (0, 0, 0), -1j * dt, normalize=True, target=1.0)
    eo = orc.gpe_energy(a, v, 0.0, step, 1.0)
    assert steps == 20
    assert rel_l2(got, a) < 1e-12
    assert abs(e.value - eo) <= 1e-12 * abs(eo)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_absorbing_boundary_every_length(lg, orc, shape):
    """grid_wf_boundary + grid_wf_propagate (2nd-order FFT and CFFT, FP64) over the same shapes:
    two fused steps kept in real space and the merge sweep, against the oracle."""
    print(f"[case] absorbing {shape}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(9 + sum(shape))
    step, mass, time = 0.3, 1.0, 0.03 - 0.005j
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1 + 1.0
    pot = (0.5 * rng.uniform(0, 1, shape) + 0.01j * rng.uniform(0, 1, shape)).astype(np.complex128)
    bounds = []
    for n in shape:
        bounds += [0, 0] if n == 1 else [n // 4, n - 1 - n // 4]
    vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
    LG.set_cgrid(vp, pot)
    for prop, km in ((LG.WF_2ND_ORDER_FFT, None), (LG.WF_2ND_ORDER_CFFT, kmax_for(shape, step, 0.6))):
        w = lg.grid_wf_alloc(*shape, step, mass, LG.WF_PERIODIC_BOUNDARY, prop, b"w")
        if km is not None:
            w.contents.kmax = km
        LG.set_cgrid(w.contents.grid, psi)
        lg.grid_wf_boundary(w, None, *bounds)
        ref = psi.copy()
        for _ in range(3):
            lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
            orc.wf_propagate_absorb_inplace(ref, pot, step, mass, (0, 0, 0), time, tuple(bounds), km)
        err = rel_l2(LG.get_cgrid(w.contents.grid), ref)
        lg.grid_wf_boundary(w, None, 0, 0, 0, 0, 0, 0)
        lg.grid_wf_free(w)
        assert err < 1e-12, f"{shape} propagator {prop} bounds {bounds}: rel L2 {err:.3e}"
    lg.cgrid_free(vp)
