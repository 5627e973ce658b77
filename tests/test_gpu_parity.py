"""GPU parity: libgrid.so's HIP path against the CPU oracle (and the committed numpy golden
vectors) on identical seeded inputs, through the C API.

Bars (FP64): relative L2 <= 1e-12 for transforms and propagation, |delta norm| <= 1e-13 per
real-time step, energies to 1e-12 relative.  Every test prints its case first so a failing
run names it.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import libgrid_amd as LG
from conftest import rel_l2
from libgrid_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")
TOL = 1e-12


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def new_cgrid(lib, arr, step=0.2, name=b"g"):
    nx, ny, nz = arr.shape
    g = lib.cgrid_alloc(nx, ny, nz, step, LG.boundary("cgrid_value_outside_periodic"), None, name)
    LG.set_cgrid(g, arr)
    return g


def new_rgrid(lib, arr, step=0.2, name=b"r"):
    nx, ny, nz = arr.shape
    g = lib.rgrid_alloc(nx, ny, nz, step, LG.boundary("rgrid_value_outside_periodic"), None, name)
    LG.set_rgrid(g, arr)
    return g


def new_wf(lib, psi, step, mass=1.0):
    nx, ny, nz = psi.shape
    w = lib.grid_wf_alloc(nx, ny, nz, step, mass, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"wf")
    LG.set_cgrid(w.contents.grid, psi)
    return w


def rand_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)


# powers of two up to 1024 and 3 * 2^k up to 384 on every axis, in 1-, 2- and 3-D shapes
FFT_SHAPES = [(16, 16, 16), (8, 4, 2), (4, 32, 8), (2, 2, 256), (1, 1, 64), (64, 64, 64), (32, 64, 128), (512, 4, 2),
              (2, 1024, 2), (2, 16, 1), (1, 1, 1), (128, 1, 1), (1, 256, 256),
              (48, 96, 192), (1, 192, 192), (2, 384, 2), (3, 6, 12), (24, 1, 96), (384, 2, 3), (96, 64, 48),
              # 512-point axes (three register stages): contiguous, strided, partial workgroups (18
              # and 36 pencils against 4 or 8 per workgroup)
              (3, 6, 512), (3, 512, 6), (512, 3, 12), (16, 512, 32), (512, 2, 512)]


def gpu_length(n):
    """Axis lengths the pencil transforms take: powers of two up to 1024, 3 * 2^k up to 384."""
    return n == 1 or (n <= 1024 and (n & (n - 1)) == 0) or (n % 3 == 0 and n <= 384 and ((n // 3) & (n // 3 - 1)) == 0)


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_fft_forward_inverse(lg, orc, shape):
    case(f"fft {shape}")
    a = rand_c(shape, 11)
    g = new_cgrid(lg, a)
    lg.cgrid_fft(g)
    fa = LG.get_cgrid(g)
    assert rel_l2(fa, orc.fft3d(a, -1)) < TOL
    lg.cgrid_inverse_fft(g)
    assert rel_l2(LG.get_cgrid(g), orc.fft3d(fa, +1)) < TOL
    lg.cgrid_fft(g)
    lg.cgrid_inverse_fft_norm(g)
    assert rel_l2(LG.get_cgrid(g), orc.inverse_fft_norm(orc.fft3d(orc.fft3d(fa, +1), -1))) < TOL
    lg.cgrid_free(g)


def test_fft_golden_vectors(lg):
    gold = np.load(GOLDEN)
    for tag in sorted(k[3:] for k in gold.files if k.startswith("in_")):
        shape = tuple(int(s) for s in tag.split("x"))
        if not all(gpu_length(n) for n in shape):
            continue  # lengths with a factor 5 or 9 (6x10x3, 1x12x9) are outside the GPU transform set
        case(f"golden {tag}")
        g = new_cgrid(lg, gold["in_" + tag])
        lg.cgrid_fft(g)
        assert rel_l2(LG.get_cgrid(g), gold["fwd_" + tag]) < TOL, tag
        LG.set_cgrid(g, gold["in_" + tag])
        lg.cgrid_inverse_fft(g)
        assert rel_l2(LG.get_cgrid(g), gold["bwd_" + tag]) < TOL, tag
        lg.cgrid_free(g)


FFT32_SHAPES = [(16, 16, 16), (3, 6, 512), (3, 512, 6), (512, 3, 12), (16, 512, 32), (1, 192, 96)]


@pytest.mark.parametrize("shape", FFT32_SHAPES)
def test_fft_single_precision(lg, orc, shape):
    """complex64 storage and FP32 butterflies with the FP64-accurate twiddle table rounded once:
    relative L2 against the FP64 oracle of the FP32-rounded input stays at FP32 rounding."""
    case(f"fft fp32 {shape}")
    a = rand_c(shape, 19).astype(np.complex64).astype(np.complex128)
    g = new_cgrid(lg, a)
    lg.cgrid_set_precision(g, b"\x01")
    lg.cgrid_fft(g)
    fa = LG.get_cgrid(g)
    err = rel_l2(fa, orc.fft3d(a, -1))
    lg.cgrid_inverse_fft(g)
    err2 = rel_l2(LG.get_cgrid(g), a * a.size)
    print(f"fp32 fft {shape}: forward {err:.2e}, round trip {err2:.2e}")
    assert err < 2e-6 and err2 < 4e-6
    lg.cgrid_free(g)


def test_fft_length_set_matches_library(lg):
    """The lengths the tests treat as supported are exactly those the library accepts."""
    lg.hgrid_fft_supported.restype = C.c_int
    lg.hgrid_fft_supported.argtypes = [C.c_long]
    for n in range(2, 1100):
        assert bool(lg.hgrid_fft_supported(n)) == gpu_length(n), n


@pytest.mark.parametrize("shape", [(8, 16, 12), (1, 48, 64), (1, 1, 96)])
def test_inverse_fft_norm2(lg, orc, shape):
    """cgrid_inverse_fft_norm2 (src/cgrid.c:1887-1890): the inverse transform scaled by
    fft_norm2 = step^d / N, d = number of axes longer than 1."""
    case(f"inverse_fft_norm2 {shape}")
    step = 0.3
    a = rand_c(shape, 17)
    g = new_cgrid(lg, a, step)
    d = sum(1 for n in shape if n > 1)
    assert abs(g.contents.fft_norm2 - step**d / a.size) < 1e-18
    lg.cgrid_inverse_fft_norm2(g)
    assert rel_l2(LG.get_cgrid(g), orc.fft3d(a, +1) * (step**d / a.size)) < TOL
    # after a forward transform it is the spatially normalised round trip
    LG.set_cgrid(g, a)
    lg.cgrid_fft(g)
    lg.cgrid_inverse_fft_norm2(g)
    assert rel_l2(LG.get_cgrid(g), a * step**d) < TOL
    lg.cgrid_free(g)


@pytest.mark.parametrize("boundary,dirichlet", [("neumann", 0j), ("dirichlet", 0j), ("dirichlet", 0.3 - 0.2j),
                                                ("periodic", 0j)])
def test_fd_energy_boundary_rules(lg, orc, boundary, dirichlet):
    """grid_wf_energy's finite-difference branch (src/grid_wf_cn.c:19-49 with the 7-point
    Laplacian of src/cgrid_diff.c:303-330) under each boundary function of src/cgrid_bc.c:21-86,
    with a pending scale on the grid, against the oracle."""
    case(f"fd energy {boundary} dirichlet={dirichlet}")
    shape, step, mass = (12, 10, 14), 0.3, 1.3
    a = rand_c(shape, 27) + 0.5
    v = np.random.default_rng(28).uniform(0, 1, shape)
    bcs = {"neumann": ("cgrid_value_outside_neumann", 1), "dirichlet": ("cgrid_value_outside_dirichlet", 0),
           "periodic": ("cgrid_value_outside_periodic", 2)}
    fname, bc = bcs[boundary]
    w = lg.grid_wf_alloc(*shape, step, mass, LG.WF_PERIODIC_BOUNDARY, LG.WF_2ND_ORDER_FFT, b"fd")
    dval = (C.c_double * 2)(dirichlet.real, dirichlet.imag)
    w.contents.grid.contents.value_outside = LG.boundary(fname)
    if dirichlet != 0:
        w.contents.grid.contents.outside_params_ptr = C.cast(dval, C.c_void_p)
    LG.set_cgrid(w.contents.grid, a)
    s = 0.8 + 0.6j
    lg.cgrid_multiply(w.contents.grid, LG.c_complex.of(s))  # pending scale: multiplies the values, not the boundary value
    vr = new_rgrid(lg, v, step)
    lg.grid_wf_analyze_method(b"\x00")
    e = lg.grid_wf_energy(w, vr)
    ref = orc.kinetic_energy_fd(a * s, step, mass, bc, dirichlet) + orc.grid_expectation(a * s, v, step)
    assert abs(e - ref) < 1e-12 * abs(ref)
    t = lg.grid_wf_kinetic_energy_cn(w)
    assert abs(t - orc.kinetic_energy_fd(a * s, step, mass, bc, dirichlet)) < 1e-12 * abs(ref)
    lg.grid_wf_free(w)
    lg.rgrid_free(vr)


def test_scaled_inverse_and_lazy_multiply(lg, orc):
    case("scaled inverse / multiply")
    a = rand_c((8, 16, 32), 3)
    g = new_cgrid(lg, a)
    lg.cgrid_multiply(g, LG.c_complex.of(0.5 - 2j))
    lg.cgrid_fft(g)
    lg.cgrid_scaled_inverse_fft(g, LG.c_complex.of(0.25 + 1j))
    expect = orc.fft3d(orc.fft3d(a * (0.5 - 2j), -1), +1) * (0.25 + 1j)
    assert rel_l2(LG.get_cgrid(g), expect) < TOL
    v = lg.cgrid_value_at_index(g, 3, 5, 7).value()
    assert abs(v - expect[3, 5, 7]) < 1e-12 * np.abs(expect).max()
    lg.cgrid_free(g)


@pytest.mark.parametrize("shape,k0,time", [((16, 32, 64), (0.0, 0.0, 0.0), 0.05), ((8, 8, 8), (0.3, -0.2, 0.1), 0.03 - 0.02j),
                                           ((1, 64, 32), (0.0, 0.5, 0.0), -0.1j),
                                           # 512 points on the outer axis: 16-pencil workgroups,
                                           # the last one partial (24 and 36 pencils)
                                           ((512, 2, 12), (0.2, 0.0, 0.0), -0.1j), ((512, 3, 12), (0.0, 0.0, 0.0), 0.03)])
def test_kinetic_phase(lg, orc, shape, k0, time):
    case(f"kinetic {shape} k0={k0} t={time}")
    step, mass = 0.3, 1.7
    a = rand_c(shape, 5)
    w = new_wf(lg, a, step, mass)
    lg.cgrid_set_momentum(w.contents.grid, *k0)
    lg.grid_wf_propagate_kinetic_fft(w, LG.c_complex.of(time))
    assert rel_l2(LG.get_cgrid(w.contents.grid), orc.propagate_kinetic_fft(a, step, mass, k0, time)) < TOL
    lg.grid_wf_free(w)


@pytest.mark.parametrize("shape", [(512, 2, 6), (16, 24, 8), (1024, 2, 2), (24, 4, 4), (6, 4, 4), (8, 4, 4),
                                   (256, 2, 4)])
@pytest.mark.parametrize("time", [-0.05j, 0.02])
def test_kinetic_phase_fp32(lg, orc, shape, time):
    """complex64 storage, real (imaginary time) and complex kinetic factors on the outer axis, at
    the lengths whose FP32 real-factor pass runs the kernel with both factor paths (6, 8, 16, 24,
    512, 1024) and at one that does not (256), against the FP64 oracle of the rounded input."""
    case(f"fp32 kinetic {shape} t={time}")
    step, mass = 0.3, 1.3
    a = rand_c(shape, 8).astype(np.complex64).astype(np.complex128)
    w = new_wf(lg, a, step, mass)
    lg.cgrid_set_precision(w.contents.grid, b"\x01")
    lg.grid_wf_propagate_kinetic_fft(w, LG.c_complex.of(time))
    err = rel_l2(LG.get_cgrid(w.contents.grid), orc.propagate_kinetic_fft(a, step, mass, (0, 0, 0), time))
    print(f"rel L2 {err:.2e}")
    assert err < 2e-7  # measured 2.5e-8 to 2.8e-8: the FP32 roundings of two transforms
    lg.grid_wf_free(w)


@pytest.mark.parametrize("tstep,cons", [(0.05, 0.0), (0.02 - 0.03j, 0.4), (-0.05j, -0.1)])
def test_potential_phase(lg, orc, tstep, cons):
    case(f"potential t={tstep} cons={cons}")
    shape = (16, 16, 32)
    a = rand_c(shape, 6)
    rng = np.random.default_rng(9)
    pot = (rng.uniform(0, 2, shape) - 0.1j * rng.uniform(0, 1, shape)).astype(np.complex128)
    w = new_wf(lg, a, 0.2)
    vp = new_cgrid(lg, pot)
    lg.grid_wf_propagate_potential(w, LG.c_complex.of(tstep), vp, cons)
    assert rel_l2(LG.get_cgrid(w.contents.grid), orc.propagate_potential(a, pot, tstep, cons)) < TOL
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)


@pytest.mark.parametrize("shape,k0,time,nullpot", [
    ((32, 32, 32), (0, 0, 0), 0.05, False),
    ((16, 32, 64), (0.0, 0.3, -0.2), 0.04 - 0.01j, False),
    ((1, 64, 128), (0, 0, 0), -0.05j, False),
    ((64, 1, 32), (0, 0, 0), 0.05, False),
    ((1, 1, 256), (0, 0, 0.2), 0.05, False),
    ((32, 16, 16), (0, 0, 0), 0.05, True),
    ((8, 512, 4), (0, 0, 0), 0.01, False),
    ((512, 4, 6), (0.0, 0.1, 0.0), 0.01, False),
    ((512, 6, 4), (0, 0, 0), -0.02j, False),
])
def test_propagate_matches_oracle(lg, orc, shape, k0, time, nullpot):
    case(f"propagate {shape} k0={k0} t={time} nullpot={nullpot}")
    step, mass = 0.2, 1.0
    rng = np.random.default_rng(4)
    psi = W.normalize(rand_c(shape, 12) * 0.1 + 1.0, step)
    pot = (0.5 * rng.uniform(0, 1, shape) + 0.01j * rng.uniform(0, 1, shape)).astype(np.complex128)
    w = new_wf(lg, psi, step, mass)
    lg.cgrid_set_momentum(w.contents.grid, *k0)
    vp = None if nullpot else new_cgrid(lg, pot)
    ref = psi.copy()
    for _ in range(3):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(time))
        orc.wf_propagate_inplace(ref, None if nullpot else pot, step, mass, k0, time)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    lg.grid_wf_free(w)
    if vp:
        lg.cgrid_free(vp)


def test_wavepacket_harmonic_128(lg, orc):
    """The 128^3 packet-in-a-trap workload: final psi, per-step norm and FFT energy vs the oracle."""
    psi0, v, prm = W.wavepacket_harmonic(seed=0)
    steps = 20
    case(f"wavepacket 128^3, {steps} of {prm['steps']} steps, params {prm}")
    step, mass, dt = prm["step"], prm["mass"], prm["dt"]
    pot = v.astype(np.complex128)
    w = new_wf(lg, psi0, step, mass)
    vp = new_cgrid(lg, pot)
    vr = new_rgrid(lg, v, step)
    ref = psi0.copy()
    n0 = lg.grid_wf_norm(w)
    for s in range(steps):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(dt))
        orc.wf_propagate_inplace(ref, pot, step, mass, (0, 0, 0), dt)
        n1 = lg.grid_wf_norm(w)
        assert abs(n1 - n0) <= 1e-13, f"norm drift at step {s}"
        n0 = n1
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    lg.grid_wf_analyze_method(b"\x01")
    e_gpu = lg.grid_wf_energy(w, vr)
    e_ref = orc.energy_fft(ref, v, step, mass)
    assert abs(e_gpu - e_ref) < 1e-12 * abs(e_ref)
    lg.grid_wf_analyze_method(b"\x00")
    e_fd = lg.grid_wf_energy(w, vr)
    e_fd_ref = orc.kinetic_energy_fd_periodic(ref, step, mass) + orc.grid_expectation(ref, v, step)
    assert abs(e_fd - e_fd_ref) < 1e-12 * abs(e_fd_ref)
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)
    lg.rgrid_free(vr)


def test_norm_normalize_overlap_expectation(lg, orc):
    case("norm / normalize / overlap / expectation / laplace expectation")
    shape, step = (16, 32, 32), 0.25
    a = rand_c(shape, 21)
    b = rand_c(shape, 22)
    v = np.random.default_rng(23).uniform(0, 1, shape)
    wa = new_wf(lg, a, step)
    wb = new_wf(lg, b, step)
    vr = new_rgrid(lg, v, step)
    assert abs(lg.grid_wf_norm(wa) - orc.integral_of_square(a, step)) < 1e-13 * orc.integral_of_square(a, step)
    ov = lg.grid_wf_overlap(wa, wb).value()
    ovr = orc.integral_of_conjugate_product(a, b, step)
    assert abs(ov - ovr) < 1e-12 * abs(ovr)
    ge = lg.grid_grid_expectation_value(wa.contents.grid, vr)
    assert abs(ge - orc.grid_expectation(a, v, step)) < 1e-12 * abs(ge)
    nrm = lg.grid_wf_normalize(wa)
    assert abs(nrm - orc.integral_of_square(a, step)) < 1e-12 * nrm
    assert abs(lg.grid_wf_norm(wa) - 1.0) < 1e-14
    an, _ = orc.wf_normalize(a, step)
    assert rel_l2(LG.get_cgrid(wa.contents.grid), an) < 1e-14
    lg.cgrid_fft(wb.contents.grid)
    le = lg.cgrid_fft_laplace_expectation_value(wb.contents.grid)
    assert abs(le - orc.laplace_expectation(orc.fft3d(b, -1), step)) < 1e-12 * abs(le)
    lg.grid_wf_free(wa)
    lg.grid_wf_free(wb)
    lg.rgrid_free(vr)


def test_density_and_real_to_complex(lg):
    case("density / real_to_complex")
    shape, step = (8, 8, 16), 0.3
    a = rand_c(shape, 31)
    w = new_wf(lg, a, step)
    r = lg.rgrid_alloc(*shape, step, LG.boundary("rgrid_value_outside_periodic"), None, b"rho")
    lg.grid_wf_density(w, r)
    rho = LG.get_rgrid(r)
    assert np.max(np.abs(rho - np.abs(a) ** 2)) < 1e-14 * np.max(np.abs(a) ** 2)
    c = lg.cgrid_alloc(*shape, step, LG.boundary("cgrid_value_outside_periodic"), None, b"c")
    lg.grid_real_to_complex_re(c, r)
    assert np.max(np.abs(LG.get_cgrid(c) - rho)) == 0.0
    lg.grid_wf_free(w)
    lg.rgrid_free(r)
    lg.cgrid_free(c)


@pytest.mark.parametrize("normalize", [True, False])
def test_gpe_step_matches_composition(lg, orc, normalize):
    case(f"gpe step normalize={normalize}")
    psi, vext, prm = W.gpe_barrier(seed=3, n=32, step=0.8)
    step, g, dt = prm["step"], prm["g"], prm["dt"]
    w = new_wf(lg, psi, step)
    vr = new_rgrid(lg, vext, step)
    ref = psi.copy()
    energies_gpu, energies_ref = [], []
    for _ in range(4):
        energies_ref.append(orc.gpe_energy(ref, vext, g, step, 1.0))
        lg.grid_wf_gpe_propagate(w, vr, g, LG.c_complex.of(-1j * dt), 1 if normalize else 0)
        energies_gpu.append(lg.grid_wf_gpe_last_energy(w))
        orc.gpe_step_inplace(ref, vext, g, step, 1.0, (0, 0, 0), -1j * dt, normalize=normalize)
    assert rel_l2(LG.get_cgrid(w.contents.grid), ref) < TOL
    for eg, er in zip(energies_gpu, energies_ref):
        assert abs(eg - er) < 1e-12 * abs(er)
    assert abs(lg.grid_wf_gpe_energy(w, vr, g) - orc.gpe_energy(ref, vext, g, step, 1.0)) < 1e-12 * abs(energies_ref[-1])
    lg.grid_wf_free(w)
    lg.rgrid_free(vr)


def test_potential_keeps_its_precision(lg, orc):
    """A complex potential used with an FP32 state keeps its FP64 values: the state's storage
    precision gets a converted copy.  Propagating an FP64 state with the same potential afterwards
    matches the oracle at the FP64 bar (had the potential been rounded to FP32 in place, the
    result would be off by ~5e-10)."""
    case("potential precision after FP32 use")
    shape, step, mass, dt = (16, 24, 32), 0.3, 1.0, 0.02
    rng = np.random.default_rng(21)
    psi = rand_c(shape, 22) * 0.1 + 1.0
    pot = (rng.uniform(0, 1, shape) + 0.05j * rng.uniform(0, 1, shape)).astype(np.complex128)
    vp = new_cgrid(lg, pot, step)
    w32 = new_wf(lg, psi, step, mass)
    lg.cgrid_set_precision(w32.contents.grid, b"\x01")
    for _ in range(2):
        lg.grid_wf_propagate(w32, vp, LG.c_complex.of(dt))
    assert lg.cgrid_get_precision(vp) == b"\x00"
    assert rel_l2(LG.get_cgrid(vp), pot) == 0.0
    w64 = new_wf(lg, psi, step, mass)
    ref = psi.copy()
    for _ in range(3):
        lg.grid_wf_propagate(w64, vp, LG.c_complex.of(dt))
        orc.wf_propagate_inplace(ref, pot, step, mass, (0, 0, 0), dt)
    assert rel_l2(LG.get_cgrid(w64.contents.grid), ref) < TOL
    # the potential written after the FP32 use is seen by the next FP32 step (the copy follows it)
    pot2 = pot * 0.5
    LG.set_cgrid(vp, pot2)
    ref32 = LG.get_cgrid(w32.contents.grid).copy()
    lg.grid_wf_propagate(w32, vp, LG.c_complex.of(dt))
    orc.wf_propagate_inplace(ref32, pot2, step, mass, (0, 0, 0), dt)
    assert rel_l2(LG.get_cgrid(w32.contents.grid), ref32) < 2e-6
    for w in (w32, w64):
        lg.grid_wf_free(w)
    lg.cgrid_free(vp)


def test_single_precision_path(lg, orc):
    """complex64 storage with FP64 twiddles / phase arguments / reductions: error vs FP64 oracle."""
    psi0, v, prm = W.wavepacket_harmonic(seed=1, n=64, step=0.4)
    case(f"fp32 path 64^3 {prm}")
    step, dt = prm["step"], prm["dt"]
    pot = v.astype(np.complex128)
    w = new_wf(lg, psi0, step)
    lg.cgrid_set_precision(w.contents.grid, b"\x01")
    vp = new_cgrid(lg, pot)
    ref = psi0.copy()
    n0 = lg.grid_wf_norm(w)
    for _ in range(20):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(dt))
        orc.wf_propagate_inplace(ref, pot, step, 1.0, (0, 0, 0), dt)
    err = rel_l2(LG.get_cgrid(w.contents.grid), ref)
    drift = abs(lg.grid_wf_norm(w) - n0)
    print(f"fp32: rel L2 {err:.3e}, norm drift {drift:.3e} after 20 steps")
    assert err < 2e-5
    assert drift < 5e-5
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)


@pytest.mark.parametrize("steps", [20, 200])
def test_single_storage_fp64_arithmetic_path(lg, orc, steps):
    """complex64 storage with FP64 arithmetic in every pass (GRID_PREC_SINGLE_STORAGE): the only
    rounding is the store, so the error against the FP64 oracle is a random walk of FP32
    rounding, far below the plain FP32 path's, and the norm does not drift systematically."""
    psi0, v, prm = W.wavepacket_harmonic(seed=1, n=64, step=0.4)
    case(f"fp32-storage fp64-arithmetic path 64^3, {steps} steps")
    step, dt = prm["step"], prm["dt"]
    pot = v.astype(np.complex128)
    w = new_wf(lg, psi0, step)
    lg.cgrid_set_precision(w.contents.grid, b"\x02")
    assert lg.cgrid_get_precision(w.contents.grid) == b"\x02"
    vp = new_cgrid(lg, pot)
    ref = psi0.copy()
    n0 = lg.grid_wf_norm(w)
    for _ in range(steps):
        lg.grid_wf_propagate(w, vp, LG.c_complex.of(dt))
    for _ in range(steps):
        orc.wf_propagate_inplace(ref, pot, step, 1.0, (0, 0, 0), dt)
    err = rel_l2(LG.get_cgrid(w.contents.grid), ref)
    drift = abs(lg.grid_wf_norm(w) - n0)
    print(f"fp32 storage / fp64 arithmetic: rel L2 {err:.3e}, norm drift {drift:.3e} after {steps} steps")
    assert err < 2e-6 * np.sqrt(steps / 20)
    assert drift < 2e-6 * np.sqrt(steps / 20)
    lg.grid_wf_free(w)
    lg.cgrid_free(vp)


def test_add_scaled_and_merge(lg):
    """cgrid_add_scaled (reference src/cgrid.c:1303-1319) against numpy, with a pending scale on
    the source, and grid_wf_merge (src/grid_wf.c:728-755) against its definition with
    grid_wf_absorb's amplitude."""
    case("cgrid_add_scaled and grid_wf_merge, 16x12x20")
    shape = (16, 12, 20)
    a, c = rand_c(shape, 31), rand_c(shape, 32)
    ga, gc = new_cgrid(lg, a), new_cgrid(lg, c)
    lg.cgrid_multiply(ga, LG.c_complex.of(0.5 - 0.25j))  # pending scale, folded in by the call
    d = 1.5 + 0.75j
    lg.cgrid_add_scaled(gc, LG.c_complex.of(d), ga)
    assert rel_l2(LG.get_cgrid(gc), c + d * (0.5 - 0.25j) * a) < 1e-15
    b = (3, 11, 2, 8, 4, 14)
    w = new_wf(lg, c, 0.2)
    lg.grid_wf_boundary(w, None, *b)
    gr, gi = new_cgrid(lg, a), new_cgrid(lg, c)
    lg.grid_wf_merge(w, gr, gi)
    alpha = np.array([[[lg.grid_wf_absorb(i, j, k, *b) for k in range(shape[2])] for j in range(shape[1])]
                      for i in range(shape[0])])
    assert np.any(alpha == 0) and np.any(alpha == 1) and np.any((alpha > 0) & (alpha < 1))
    assert rel_l2(LG.get_cgrid(w.contents.grid), (1 - alpha) * a + alpha * c) < 1e-15
    for g in (ga, gc, gr, gi):
        lg.cgrid_free(g)
    lg.grid_wf_free(w)
