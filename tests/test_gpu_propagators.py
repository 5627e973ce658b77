"""GPU parity of the propagator variants against the CPU oracle on the same inputs:

  * WF_2ND_ORDER_CFFT (reference src/grid_wf.c:491-498 with grid_wf_propagate_kinetic_cfft,
    src/grid_wf_fft.c:206-262): the split step whose kinetic phase removes the components
    with |k - k0| >= kmax.  The cutoff test sums (k - k0)^2 in the reference's order x, y, z,
    so the set of removed components is the oracle's exactly.  kmax is chosen so that a
    sizeable share of the spectrum is removed.
  * WF_4TH_ORDER_FFT / WF_4TH_ORDER_CFFT (src/grid_wf.c:433-457, :505-525):
    V(t/6) K(t/2) V(2t/3) K(t/2) V(t/6);
  * the absorbing boundary of the 2nd-order ones (grid_wf_boundary, src/grid_wf.c:396-422):
    a real-time and an imaginary-time step merged by the boundary amplitude.

The bar is the FP64 one of the plain step (relative L2 <= 1e-12).
"""
import sys

import numpy as np
import pytest

import libgrid_amd as LG
from conftest import rel_l2
from libgrid_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rand_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)


def new_wf(lib, psi, step, mass, kmax, propagator=LG.WF_2ND_ORDER_CFFT):
    nx, ny, nz = psi.shape
    w = lib.grid_wf_alloc(nx, ny, nz, step, mass, LG.WF_PERIODIC_BOUNDARY, propagator, b"wf")
    if kmax is not None:
        w.contents.kmax = kmax
    LG.set_cgrid(w.contents.grid, psi)
    return w


def kmax_for(shape, step, frac):
    """frac times the largest |k| on the grid (Nyquist along every non-trivial axis)."""
    kk = sum((np.pi / step) ** 2 for n in shape if n > 1)
    return frac * np.sqrt(kk)


@pytest.mark.parametrize("shape,k0,time", [((16, 32, 64), (0.0, 0.0, 0.0), 0.05), ((8, 8, 8), (0.3, -0.2, 0.1), 0.03 - 0.02j),
                                           ((1, 64, 32), (0.0, 0.5, 0.0), -0.1j), ((1, 1, 128), (0.0, 0.0, 0.3), 0.02)])
def test_kinetic_cfft(lg, orc, shape, k0, time):
    case(f"kinetic cfft {shape} k0={k0} t={time}")
    step, mass = 0.3, 1.7
    kmax = kmax_for(shape, step, 0.6)
    a = rand_c(shape, 5)
    w = new_wf(lg, a, step, mass, kmax)
    lg.cgrid_set_momentum(w.contents.grid, *k0)
    lg.grid_wf_propagate_kinetic_cfft(w, LG.c_complex.of(time))
    ref = orc.propagate_kinetic_cfft(a, step, mass, k0, time, kmax)
    out = LG.get_cgrid(w.contents.grid)
    assert np.array_equal(out == 0, ref == 0), "cut sets differ"
    assert 0.2 < np.mean(ref == 0) < 0.9
    assert rel_l2(out, ref) < TOL
    lg.grid_wf_free(w)


@pytest.mark.parametrize("shape,k0,time,frac,nullpot", [
    ((32, 32, 32), (0, 0, 0), 0.05, 0.5, False),
    ((16, 32, 64), (0.0, 0.3, -0.2), 0.04 - 0.01j, 0.7, False),
    ((32, 32, 32), (0.2, 0.0, 0.1), -0.05j, 0.4, False),
    ((1, 64, 128), (0, 0.2, 0), -0.05j, 0.6, False),
    ((64, 1, 32), (0, 0, 0), 0.05, 0.6, False),
    ((1, 1, 256), (0, 0, 0.2), 0.05, 0.5, False),
    ((32, 16, 16), (0, 0, 0), 0.05, 0.5, True),
])
def test_cfft_propagate_matches_oracle(lg, orc, shape, k0, time, frac, nullpot):
    case(f"cfft propagate {shape} k0={k0} t={time} kmax={frac}*kmax")
    step, mass = 0.2, 1.0
    kmax = kmax_for(shape, step, frac)
    rng = np.random.default_rng(4)
    psi = W.normalize(rand_c(shape, 12) * 0.1 + 1.0, step)
    pot = (0.5 * rng.uniform(0, 1, shape) + 0.01j * rng.uniform(0, 1, shape)).astype(np.complex128)
    w = new_wf(lg, psi, step, mass, kmax)
    lg.cgrid_set_momentum(w.contents.grid, *k0)
    vp = None
    if not nullpot:
        vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
        LG.set_cgrid(vp, pot)
    ref = psi.copy()
    for _ in range(3):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        orc.wf_propagate_cfft_inplace(ref, None if nullpot else pot, step, mass, k0, time, kmax)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    # the cutoff did remove something: the plain step differs
    plain = psi.copy()
    for _ in range(3):
        orc.wf_propagate_inplace(plain, None if nullpot else pot, step, mass, k0, time)
    assert rel_l2(plain, ref) > 1e-6
    lg.grid_wf_free(w)
    if vp:
        lg.cgrid_free(vp)


def test_cfft_gpe_step_matches_composition(lg, orc):
    """The fused Gross-Pitaevskii step with the CFFT propagator (imaginary time, renormalised)
    against the oracle's composition of reference calls with the cut-off kinetic phase."""
    case("gpe cfft step")
    psi, vext, prm = W.gpe_barrier(seed=3, n=32, step=0.8)
    step, g, dt = prm["step"], prm["g"], prm["dt"]
    kmax = kmax_for(psi.shape, step, 0.5)
    w = new_wf(lg, psi, step, 1.0, kmax)
    vr = lg.rgrid_alloc(*psi.shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"v")
    LG.set_rgrid(vr, vext)
    ref = psi.copy()
    for _ in range(4):
        lg.grid_wf_gpe_propagate(w, vr, g, LG.c_complex.of(-1j * dt), 1)
        orc.gpe_step_inplace(ref, vext, g, step, 1.0, (0, 0, 0), -1j * dt, normalize=True, kmax=kmax)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    assert abs(lg.grid_wf_gpe_energy(w, vr, g) - orc.gpe_energy(ref, vext, g, step, 1.0)) < 1e-12 * abs(
        orc.gpe_energy(ref, vext, g, step, 1.0))
    lg.grid_wf_free(w)
    lg.rgrid_free(vr)


@pytest.mark.parametrize("shape,k0,time,frac,nullpot", [
    ((32, 32, 32), (0, 0, 0), 0.05, None, False),
    ((16, 32, 64), (0.0, 0.3, -0.2), 0.04 - 0.01j, None, False),
    ((32, 32, 32), (0, 0, 0), -0.05j, None, False),
    ((1, 64, 128), (0, 0.2, 0), 0.05, None, False),
    ((1, 1, 256), (0, 0, 0.2), 0.05, None, False),
    ((32, 16, 16), (0, 0, 0), 0.05, None, True),
    ((32, 32, 32), (0.2, 0.0, 0.1), 0.05, 0.5, False),
    ((1, 1, 256), (0, 0, 0), -0.05j, 0.5, False),
])
def test_fourth_order_matches_oracle(lg, orc, shape, k0, time, frac, nullpot):
    case(f"4th order {shape} k0={k0} t={time} cutoff={frac} nullpot={nullpot}")
    step, mass = 0.2, 1.0
    kmax = None if frac is None else kmax_for(shape, step, frac)
    prop = LG.WF_4TH_ORDER_FFT if frac is None else LG.WF_4TH_ORDER_CFFT
    rng = np.random.default_rng(4)
    psi = W.normalize(rand_c(shape, 12) * 0.1 + 1.0, step)
    pot = (0.5 * rng.uniform(0, 1, shape) + 0.01j * rng.uniform(0, 1, shape)).astype(np.complex128)
    w = new_wf(lg, psi, step, mass, kmax, prop)
    lg.cgrid_set_momentum(w.contents.grid, *k0)
    vp = None
    if not nullpot:
        vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
        LG.set_cgrid(vp, pot)
    ref = psi.copy()
    for _ in range(3):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        orc.wf_propagate_4th_inplace(ref, None if nullpot else pot, step, mass, k0, time, kmax)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    lg.grid_wf_free(w)
    if vp:
        lg.cgrid_free(vp)


def test_fourth_order_after_second_order_steps(lg, orc):
    """A 3-D FP64 state left transformed along z by 2nd-order steps is brought back to real
    space before a 4th-order step (the propagator field switched between calls)."""
    case("2nd then 4th order, 32^3")
    shape, step, mass, time = (32, 32, 32), 0.2, 1.0, 0.05
    rng = np.random.default_rng(5)
    psi = W.normalize(rand_c(shape, 13) * 0.1 + 1.0, step)
    pot = (0.5 * rng.uniform(0, 1, shape) + 0j).astype(np.complex128)
    w = new_wf(lg, psi, step, mass, None, LG.WF_2ND_ORDER_FFT)
    vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
    LG.set_cgrid(vp, pot)
    ref = psi.copy()
    for _ in range(2):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        orc.wf_propagate_inplace(ref, pot, step, mass, (0, 0, 0), time)
    w.contents.propagator = LG.WF_4TH_ORDER_FFT
    for _ in range(2):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        orc.wf_propagate_4th_inplace(ref, pot, step, mass, (0, 0, 0), time)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)


@pytest.mark.parametrize("shape,k0,time,frac,bounds", [
    ((32, 32, 32), (0, 0, 0), 0.05, None, (6, 25, 5, 26, 4, 27)),
    ((16, 32, 64), (0.0, 0.3, -0.2), 0.04 - 0.01j, None, (3, 12, 6, 25, 10, 53)),
    ((32, 32, 32), (0.1, 0, 0), 0.05, 0.6, (6, 25, 5, 26, 4, 27)),
    ((1, 64, 64), (0, 0, 0), 0.05, None, (0, 0, 8, 55, 8, 55)),
    ((1, 1, 256), (0, 0, 0.2), 0.05, 0.5, (0, 0, 0, 0, 32, 223)),
])
def test_absorbing_boundary_matches_oracle(lg, orc, shape, k0, time, frac, bounds):
    """grid_wf_boundary + grid_wf_propagate (reference src/grid_wf.c:396-422, :464-490)."""
    case(f"absorbing {shape} k0={k0} t={time} cutoff={frac} bounds={bounds}")
    step, mass = 0.2, 1.0
    kmax = None if frac is None else kmax_for(shape, step, frac)
    prop = LG.WF_2ND_ORDER_FFT if frac is None else LG.WF_2ND_ORDER_CFFT
    rng = np.random.default_rng(6)
    psi = W.normalize(rand_c(shape, 14) * 0.1 + 1.0, step)
    pot = (0.5 * rng.uniform(0, 1, shape) + 0.01j * rng.uniform(0, 1, shape)).astype(np.complex128)
    w = new_wf(lg, psi, step, mass, kmax, prop)
    lg.cgrid_set_momentum(w.contents.grid, *k0)
    lg.grid_wf_boundary(w, None, *bounds)
    vp = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"pot")
    LG.set_cgrid(vp, pot)
    ref = psi.copy()
    for _ in range(3):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        orc.wf_propagate_absorb_inplace(ref, pot, step, mass, k0, time, bounds, kmax)
    out = LG.get_cgrid(w.contents.grid)
    assert rel_l2(out, ref) < TOL
    # the boundary did act: the plain step differs
    plain = psi.copy()
    for _ in range(3):
        orc.wf_propagate_inplace(plain, pot, step, mass, k0, time)
    assert rel_l2(plain, ref) > 1e-6
    # pure imaginary time skips the boundary
    lg.grid_wf_propagate(w, vp, LG.c_complex.of(-0.05j))
    orc.wf_propagate_inplace(ref, pot, step, mass, k0, -0.05j) if kmax is None else \
        orc.wf_propagate_cfft_inplace(ref, pot, step, mass, k0, -0.05j, kmax)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    lg.grid_wf_boundary(w, None, 0, 0, 0, 0, 0, 0)
    assert not w.contents.ts_func
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)
