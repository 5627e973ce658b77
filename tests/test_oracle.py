"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

The reference ships no tests or golden vectors for this path (its examples only print values),
so each oracle function is checked against something it does not contain:
  * FFT: numpy.fft (pocketfft) golden vectors committed in tests/golden, brute-force O(N^2)
    DFTs, Parseval / round trip / shift theorem / plane waves, and the cosine round trip of
    the reference's examples/cfft.c;
  * kinetic phase: closed-form free-particle evolution of a Gaussian packet (also in a moving
    frame, k0 != 0) and the reference's Nyquist convention; the cut-off (CFFT) variant against
    its limits (no cut, everything cut), the Nyquist convention of its test and plane-wave
    eigenstates inside and outside the cutoff sphere;
  * potential phase and the full step: constant potential = global phase, absorbing
    potential = decay, Ehrenfest dynamics in a harmonic trap, time reversibility, norm
    conservation, the harmonic-oscillator ground state E0 = 3/2 hbar omega;
  * norm and energies: analytic Gaussian integrals; FD energy converges as O(h^2);
  * GPE composition: uniform condensates (closed form) and the g = 0 limit;
  * the 4th-order step: equal to the 2nd-order step without a potential, a constant potential
    as a global phase, reversibility, and the closed-form imaginary-time fixed point of its
    stage sequence in a harmonic trap (which also pins the 2nd-order sequence);
  * the absorbing boundary: hand-computed amplitudes, the real-time / imaginary-time regions
    of the merged step, and the absorption of an outgoing packet.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2
from libgrid_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")


def brute_dft(a, sign):
    out = a.astype(np.complex128)
    for axis, n in enumerate(a.shape):
        k = np.arange(n)
        m = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(m, np.moveaxis(out, axis, 0), axes=(1, 0)), 0, axis)
    return out


def test_fft_matches_numpy_golden(orc):
    g = np.load(GOLDEN)
    tags = sorted(k[3:] for k in g.files if k.startswith("in_"))
    assert len(tags) >= 6
    for tag in tags:
        a = g["in_" + tag]
        fwd = orc.fft3d(a, -1)
        bwd = orc.fft3d(a, +1)
        assert rel_l2(fwd, g["fwd_" + tag]) < 1e-14, tag
        assert rel_l2(bwd, g["bwd_" + tag]) < 1e-14, tag


@pytest.mark.parametrize("shape", [(4, 4, 4), (2, 3, 5), (1, 1, 7), (3, 1, 8), (1, 6, 1)])
def test_fft_brute_force(orc, shape):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel_l2(orc.fft3d(a, -1), brute_dft(a, -1)) < 1e-14
    assert rel_l2(orc.fft3d(a, +1), brute_dft(a, +1)) < 1e-14


def test_fft_invariants(orc):
    rng = np.random.default_rng(3)
    shape = (8, 16, 32)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n = a.size
    fa = orc.fft3d(a, -1)
    # Parseval
    assert abs(np.sum(np.abs(fa) ** 2) / (n * np.sum(np.abs(a) ** 2)) - 1) < 1e-14
    # round trip through the normalised inverse
    assert rel_l2(orc.inverse_fft_norm(fa), a) < 1e-14
    # linearity
    assert rel_l2(orc.fft3d(2.0 * a - 3j * b, -1), 2.0 * fa - 3j * orc.fft3d(b, -1)) < 1e-14
    # shift theorem: a(x - 1) along z -> multiply by exp(-2 pi i k / nz)
    sa = np.roll(a, 1, axis=2)
    kz = np.arange(shape[2])
    assert rel_l2(orc.fft3d(sa, -1), fa * np.exp(-2j * np.pi * kz / shape[2])[None, None, :]) < 1e-13
    # a plane wave maps to N * delta at its wave vector
    i, j, k = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    pw = np.exp(2j * np.pi * (3 * i / shape[0] + 5 * j / shape[1] + 7 * k / shape[2]))
    fp = orc.fft3d(pw, -1)
    expect = np.zeros(shape, complex)
    expect[3, 5, 7] = n
    assert np.max(np.abs(fp - expect)) < 1e-9


def test_fft_cosine_round_trip_known_answer(orc):
    """examples/cfft.c:32-65: cos standing waves -> FFT -> normalised inverse FFT = input."""
    nx, ny, nz, step = 16, 32, 64, 0.5
    x, y, z = W.coords((nx, ny, nz), step)
    f = (np.cos(x * 2 * np.pi / (nx * step)) + np.cos(y * 2 * np.pi / (ny * step)) + np.cos(z * 2 * np.pi / (nz * step)))
    f = np.broadcast_to(f, (nx, ny, nz)).astype(np.complex128).copy()
    fk = orc.fft3d(f, -1)
    n = f.size
    # each cosine contributes N/2 at k = +-1 along its axis (with the sign of cos(-pi) = -1 at i = -n/2)
    mask = np.abs(fk) > 1e-6 * n
    assert mask.sum() == 6
    assert abs(np.sum(np.abs(fk[mask])) - 3 * n) < 1e-8 * n
    assert rel_l2(orc.inverse_fft_norm(fk), f) < 1e-14


def free_gaussian_1d(x, a, p, t, mass=1.0, hbar=1.0):
    """Free evolution of exp(-a x^2 + i p x): (1 + 2 i a hbar t / m)^(-1/2) exp((-a x^2 + i p x - i hbar p^2 t / 2m) / D)."""
    d = 1.0 + 2j * a * hbar * t / mass
    return d**-0.5 * np.exp((-a * x * x + 1j * p * x - 1j * hbar * p * p * t / (2 * mass)) / d)


def test_free_particle_closed_form(orc):
    """Kinetic phase and FFTs: V = 0 split-operator steps equal the exact free evolution."""
    n, step, w, mass = 128, 0.2, 1.5, 1.3
    p = np.array([0.7, -0.4, 0.25])
    shape = (n, n, n)
    a = 1.0 / w**2
    x, y, z = W.coords(shape, step)
    psi0 = (np.exp(-a * x * x + 1j * p[0] * x) * np.exp(-a * y * y + 1j * p[1] * y) * np.exp(-a * z * z + 1j * p[2] * z))
    psi = np.ascontiguousarray(psi0, dtype=np.complex128)
    dt, steps = 0.05, 10
    for _ in range(steps):
        orc.wf_propagate_inplace(psi, None, step, mass, (0.0, 0.0, 0.0), dt)
    t = dt * steps
    exact = free_gaussian_1d(x, a, p[0], t, mass) * free_gaussian_1d(y, a, p[1], t, mass) * free_gaussian_1d(z, a, p[2], t, mass)
    assert rel_l2(psi, exact) < 1e-12


def test_moving_frame_closed_form(orc):
    """k0 != 0: the reference propagates with (k - k0)^2 and adds cons = -hbar^2 k0^2 / 2m to V.
    With V = 0 the exact result is exp(-i cons t) exp(i k0 x) * free[p - k0]."""
    n, step, w, mass = 128, 0.2, 1.5, 1.0
    shape = (1, n, n)
    a = 1.0 / w**2
    l = 2 * np.pi / (n * step)
    k0 = (0.0, 2 * l, -3 * l)  # multiples of 2 pi / L as the reference requires (src/cgrid.c:290-300)
    p = (0.0, 0.5, 0.3)
    _, y, z = W.coords(shape, step)
    psi = np.ascontiguousarray(np.exp(-a * y * y + 1j * p[1] * y) * np.exp(-a * z * z + 1j * p[2] * z) * np.ones(shape))
    zero = np.zeros(shape, complex)
    dt, steps = 0.05, 10
    for _ in range(steps):
        orc.wf_propagate_inplace(psi, zero, step, mass, k0, dt)
    t = dt * steps
    cons = -(k0[1] ** 2 + k0[2] ** 2) / (2 * mass)
    exact = (np.exp(-1j * cons * t) * np.exp(1j * (k0[1] * y + k0[2] * z)) * free_gaussian_1d(y, a, p[1] - k0[1], t, mass) *
             free_gaussian_1d(z, a, p[2] - k0[2], t, mass))
    assert rel_l2(psi, np.broadcast_to(exact, shape)) < 1e-12


def test_kinetic_nyquist_convention(orc):
    """Index n/2 is the positive frequency +n/2 (src/grid_wf_fft.c:173-187); visible when k0 != 0."""
    n, step, mass = 16, 0.3, 1.0
    k0 = (0.0, 0.0, 0.4)
    ak = np.zeros((1, 1, n), complex)
    ak[0, 0, n // 2] = 1.0
    out = orc.propagate_kinetic_fft(ak, step, mass, k0, 0.1)
    kn = (n // 2) * 2 * np.pi / (n * step) - k0[2]
    assert abs(out[0, 0, n // 2] - np.exp(-1j * 0.1 * kn * kn / (2 * mass))) < 1e-14


def test_cfft_cutoff_limits_and_nyquist(orc):
    """The cut-off kinetic phase (src/grid_wf_fft.c:206-262): with a cutoff above every |k - k0|
    it is the plain kinetic phase, with kmax = 0 it zeroes everything, and the Nyquist index
    is +n/2 in the test tot < kmax^2 as well."""
    rng = np.random.default_rng(21)
    shape, step, mass, k0, t = (8, 12, 16), 0.25, 1.3, (0.1, -0.2, 0.4), 0.07 - 0.02j
    ak = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    plain = orc.propagate_kinetic_fft(ak, step, mass, k0, t)
    np.testing.assert_array_equal(orc.propagate_kinetic_cfft(ak, step, mass, k0, t, 1e4), plain)
    assert not np.any(orc.propagate_kinetic_cfft(ak, step, mass, k0, t, 0.0))
    n = 16
    one = np.zeros((1, 1, n), complex)
    one[0, 0, n // 2] = 1.0
    l = 2 * np.pi / (n * step)
    kn_pos, kn_neg = (n // 2) * l - k0[2], -(n // 2) * l - k0[2]
    kmax = np.sqrt(0.5 * ((kn_pos**2 + k0[0] ** 2 + k0[1] ** 2) + (kn_neg**2 + k0[0] ** 2 + k0[1] ** 2)))
    out = orc.propagate_kinetic_cfft(one, step, mass, k0, 0.1, kmax)
    tot = k0[0] ** 2 + k0[1] ** 2 + kn_pos**2
    assert abs(out[0, 0, n // 2] - np.exp(-1j * 0.1 * tot / (2 * mass))) < 1e-14


def test_cfft_step_plane_waves_closed_form(orc):
    """Free plane waves are eigenstates of the split step: the CFFT step keeps each wave with
    |k| < kmax, multiplied by exp(-i hbar k^2 t / 2m), and removes the others.  The set holds a
    wave whose components are each below kmax while |k| is above it (the test is on |k|^2, not
    per axis), and waves just inside and just outside the sphere."""
    shape, step, mass, t = (16, 16, 16), 0.4, 1.0, 0.3
    x, y, z = W.coords(shape, step)
    l = 2 * np.pi / (16 * step)
    modes = [((1, 0, 0), 1.0), ((2, -3, 1), 0.5 - 0.2j), ((-3, 3, 3), 0.3j), ((0, 4, 0), 0.7), ((-4, 0, 1), -0.4)]
    kmax = np.sqrt(16.5) * l  # |(-4,0,1)| = sqrt(17) l is cut, |(0,4,0)| = 4 l is kept
    psi = np.zeros(shape, complex)
    exact = np.zeros(shape, complex)
    for m, c in modes:
        k = np.array(m) * l
        wave = c * np.exp(1j * (k[0] * x + k[1] * y + k[2] * z))
        psi += wave
        if k @ k < kmax**2:
            exact += wave * np.exp(-1j * t * (k @ k) / (2 * mass))
    kept = [m for m, _ in modes if np.dot(m, m) * l * l < kmax**2]
    assert kept == [(1, 0, 0), (2, -3, 1), (0, 4, 0)]
    psi = np.ascontiguousarray(psi)
    orc.wf_propagate_cfft_inplace(psi, None, step, mass, (0.0, 0.0, 0.0), t, kmax)
    assert rel_l2(psi, exact) < 1e-13


def gaussian_fixed_point(a, b, dt, mass, kappa):
    """Width of the Gaussian exp(-alpha r^2 / 2) that one imaginary-time step
    V(a dt) K(dt/2) V(b dt) K(dt/2) V(a dt) maps onto itself in V = kappa r^2 / 2 (hbar = 1):
    V(s) adds kappa s to alpha and K(s) maps 1/alpha to 1/alpha + s/mass, exactly, because
    both act on a Gaussian in closed form.  Iterated to convergence in plain floats."""
    al = 1.0
    for _ in range(2000):
        al = 1.0 / (1.0 / (al + kappa * a * dt) + 0.5 * dt / mass)
        al = 1.0 / (1.0 / (al + kappa * b * dt) + 0.5 * dt / mass) + kappa * a * dt
    return al


@pytest.mark.parametrize("order", ["2nd", "4th"])
def test_splitting_fixed_point_closed_form(orc, order):
    """Imaginary-time steps in a harmonic trap converge to the splitting's own Gaussian (not to
    the exact ground state: the difference is O(dt^2) and depends on the stage coefficients).
    2nd order is K(t/2) V(t) K(t/2) (a = 0, b = 1); the reference's 4th-order branch is
    V(t/6) K(t/2) V(2t/3) K(t/2) V(t/6) without the |grad V|^2 term, which it forms but does
    not use (src/grid_wf.c:449-452).  At dt = 0.5 the widths of a = 1/6, 1/4 and 0 differ by
    about 1 %, so a wrong stage coefficient fails here."""
    n, step, mass, kappa, dt = 48, 0.35, 1.3, 1.0, 0.5
    a, b = (0.0, 1.0) if order == "2nd" else (1.0 / 6.0, 2.0 / 3.0)
    shape = (n, n, n)
    x, y, z = W.coords(shape, step)
    r2 = np.broadcast_to(x * x + y * y + z * z, shape)
    pot = np.ascontiguousarray(0.5 * kappa * r2 + 0j)
    psi = np.ascontiguousarray(np.exp(-0.5 * r2) + 0j)
    for _ in range(60):
        if order == "2nd":
            orc.wf_propagate_inplace(psi, pot, step, mass, (0, 0, 0), -1j * dt)
        else:
            orc.wf_propagate_4th_inplace(psi, pot, step, mass, (0, 0, 0), -1j * dt)
        psi, _ = orc.wf_normalize(psi, step)
    al = gaussian_fixed_point(a, b, dt, mass, kappa)
    exact = np.exp(-0.5 * al * r2)
    exact *= np.sqrt(1.0 / (np.sum(exact**2) * step**3))
    assert rel_l2(psi, exact) < 1e-10
    e = orc.energy_fft(psi, 0.5 * kappa * np.ascontiguousarray(r2), step, mass)
    assert abs(e - 3.0 * (al / (4.0 * mass) + kappa / (4.0 * al))) < 1e-10


def test_4th_order_limits_and_reversibility(orc):
    """Without a potential the 4th-order step is the 2nd-order one (the same FFTs and kinetic
    phases in the same order, bit for bit); a constant potential V0 is the global phase
    exp(-i (V0 + cons) t) on top of free motion; the symmetric stage sequence is reversible."""
    rng = np.random.default_rng(8)
    shape, step, mass = (16, 12, 8), 0.3, 1.1
    psi = np.ascontiguousarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    a4, a2 = psi.copy(), psi.copy()
    orc.wf_propagate_4th_inplace(a4, None, step, mass, (0.2, 0, -0.1), 0.05)
    orc.wf_propagate_inplace(a2, None, step, mass, (0.2, 0, -0.1), 0.05)
    np.testing.assert_array_equal(a4, a2)
    v0 = 0.7 - 0.05j
    a4, a2 = psi.copy(), psi.copy()
    orc.wf_propagate_4th_inplace(a4, np.full(shape, v0), step, mass, (0, 0, 0), 0.05)
    orc.wf_propagate_inplace(a2, None, step, mass, (0, 0, 0), 0.05)
    assert rel_l2(a4, a2 * np.exp(-1j * v0 * 0.05)) < 1e-13
    v = np.ascontiguousarray(rng.uniform(0, 2, shape) + 0j)
    b = psi.copy()
    orc.wf_propagate_4th_inplace(b, v, step, mass, (0, 0, 0), 0.07)
    assert rel_l2(b, psi) > 1e-3
    orc.wf_propagate_4th_inplace(b, v, step, mass, (0, 0, 0), -0.07)
    assert rel_l2(b, psi) < 1e-13


def test_absorb_amplitude_known_values(orc):
    """grid_wf_absorb (src/grid_wf.c:185-203) at hand-computed points: 0 inside the box
    [lx, hx] x [ly, hy] x [lz, hz]; outside, (distance / lower bound) per axis, summed, times
    2/3, capped at 1.  The upper side divides by the lower bound too (lx, not nx - 1 - hx)."""
    b = (4, 11, 2, 5, 3, 12)
    cases = {(6, 3, 7): 0.0, (4, 2, 3): 0.0, (11, 5, 12): 0.0, (0, 3, 7): 2 / 3, (2, 3, 7): 1 / 3,
             (13, 3, 7): 1 / 3, (15, 3, 7): 2 / 3, (6, 0, 7): 2 / 3, (6, 7, 7): 2 / 3, (6, 3, 0): 2 / 3,
             (6, 3, 14): 4 / 9, (2, 6, 7): (0.5 + 0.5) * 2 / 3, (0, 0, 0): 1.0, (15, 7, 15): 1.0}
    for (i, j, k), want in cases.items():
        assert abs(orc.wf_absorb(i, j, k, b) - want) < 1e-15, (i, j, k)


def test_absorbing_step_regions_and_absorption(orc):
    """With an absorbing boundary the step (src/grid_wf.c:396-422) is the real-time step where
    the amplitude is 0 and the imaginary-time step (time -i Re t) where it is 1, bit for bit; a
    packet that runs into the absorbing layer loses its norm, while without the boundary the
    same packet keeps it."""
    rng = np.random.default_rng(3)
    shape, step, mass, t = (16, 8, 16), 0.3, 1.0, 0.05
    b = (4, 11, 2, 5, 3, 12)
    psi = np.ascontiguousarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    pot = np.ascontiguousarray(rng.uniform(0, 1, shape) + 0j)
    ab, re_, im_ = psi.copy(), psi.copy(), psi.copy()
    orc.wf_propagate_absorb_inplace(ab, pot, step, mass, (0, 0, 0), t, b)
    orc.wf_propagate_inplace(re_, pot, step, mass, (0, 0, 0), t)
    orc.wf_propagate_inplace(im_, pot, step, mass, (0, 0, 0), -1j * t)
    alpha = np.array([[[orc.wf_absorb(i, j, k, b) for k in range(shape[2])] for j in range(shape[1])]
                      for i in range(shape[0])])
    assert np.array_equal(ab[alpha == 0], re_[alpha == 0]) and np.any(alpha == 0)
    assert np.array_equal(ab[alpha == 1], im_[alpha == 1]) and np.any(alpha == 1)
    mid = (alpha > 0) & (alpha < 1)
    assert rel_l2(ab[mid], ((1 - alpha) * re_ + alpha * im_)[mid]) < 1e-15
    # a packet with momentum +3 along z runs into the layer of indices k > 191 (and k < 64)
    n, step, mass, p = 256, 0.1, 1.0, 3.0
    _, _, z = W.coords((1, 1, n), step)
    psi = W.normalize(np.ascontiguousarray(np.exp(-((z - 2.0) ** 2) + 1j * p * z) * np.ones((1, 1, n))), step)
    free, ab = psi.copy(), psi.copy()
    bz = (0, 0, 0, 0, 64, 191)
    for _ in range(100):  # the centre moves p t / m = 9 bohr, from 4.3 before the layer into it
        orc.wf_propagate_absorb_inplace(ab, None, step, mass, (0, 0, 0), 0.03, bz)
        orc.wf_propagate_inplace(free, None, step, mass, (0, 0, 0), 0.03)
    assert abs(orc.integral_of_square(free, step) - 1.0) < 1e-12
    assert orc.integral_of_square(ab, step) < 0.25


def test_constant_and_absorbing_potential(orc):
    """A constant potential V0 is a global phase exp(-i V0 t); V0 - i G also decays as exp(-G t)."""
    rng = np.random.default_rng(5)
    shape = (8, 8, 16)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    v0, gam, dt = 0.37, 0.2, 0.05
    pot = np.full(shape, v0 - 1j * gam)
    ref_free = orc.wf_propagate(psi, None, 0.2, 1.0, (0, 0, 0), dt)
    out = orc.wf_propagate(psi, pot, 0.2, 1.0, (0, 0, 0), dt)
    assert rel_l2(out, ref_free * np.exp(-1j * v0 * dt) * np.exp(-gam * dt)) < 1e-13
    # imaginary time: exp(-tau V0)
    ref_free_it = orc.wf_propagate(psi, None, 0.2, 1.0, (0, 0, 0), -1j * dt)
    out_it = orc.wf_propagate(psi, np.full(shape, v0 + 0j), 0.2, 1.0, (0, 0, 0), -1j * dt)
    assert rel_l2(out_it, ref_free_it * np.exp(-dt * v0)) < 1e-13


def test_norm_and_energies_analytic(orc):
    n, step, w, mass = 128, 0.2, 1.5, 1.0
    rc = np.array([0.6, -1.1, 0.3])
    p = np.array([0.4, 0.9, -0.5])
    shape = (n, n, n)
    psi = W.gaussian_packet(shape, step, w, rc, p)
    assert abs(orc.integral_of_square(psi, step) / (np.pi * w * w / 2) ** 1.5 - 1) < 1e-13
    psi = W.normalize(psi, step)
    t = orc.kinetic_energy_fft(psi, step, mass)
    assert abs(t - (np.dot(p, p) + 3 / w**2) / (2 * mass)) < 1e-12
    v = W.harmonic(shape, step)
    ev = orc.grid_expectation(psi, v, step)
    assert abs(ev - 0.5 * 0.01 * (3 * w * w / 4 + np.dot(rc, rc))) < 1e-12
    assert abs(orc.energy_fft(psi, v, step, mass) - (t + ev)) < 1e-13
    # the FFT laplace expectation value of the transformed state is -|k|^2 weighted, as in the reference
    assert abs(-orc.laplace_expectation(orc.fft3d(psi, -1), step) / (2 * mass) - t) < 1e-12


def test_fd_kinetic_energy_converges(orc):
    """The 7-point FD energy (src/cgrid_diff.c:303-330) approaches the analytic value as O(h^2)."""
    w, mass, p = 1.5, 1.0, np.array([0.3, 0.0, -0.2])
    exact = (np.dot(p, p) + 3 / w**2) / (2 * mass)
    errs = []
    for n, step in ((64, 0.4), (128, 0.2)):
        psi = W.normalize(W.gaussian_packet((n, n, n), step, w, (0, 0, 0), p), step)
        errs.append(abs(orc.kinetic_energy_fd_periodic(psi, step, mass) - exact))
    assert errs[1] < 1e-2 * exact
    assert 3.0 < errs[0] / errs[1] < 5.0  # second order


def test_ehrenfest_harmonic(orc):
    """<x>(t) = x_c cos(wt) + p/(m w) sin(wt) for V = 1/2 m w^2 x^2, up to the O(dt^2) splitting error."""
    n, step, mass, omega = 128, 0.4, 1.0, 0.1
    shape = (1, 1, n)
    x = W.axes(n, step)
    xc, p = 1.2, 0.8
    w2 = 2.0 / (mass * omega)  # coherent-state width: the packet does not breathe or reach the box edge
    psi = np.ascontiguousarray(W.normalize(np.exp(-(x - xc) ** 2 / w2 + 1j * p * (x - xc)).reshape(shape), step))
    v = (0.5 * mass * omega**2 * x * x).reshape(shape).astype(complex)
    dt, steps = 0.05, 200
    for _ in range(steps):
        orc.wf_propagate_inplace(psi, v, step, mass, (0, 0, 0), dt)
    t = dt * steps
    xm = np.sum(np.abs(psi.ravel()) ** 2 * x) * step
    assert abs(xm - (xc * np.cos(omega * t) + p / (mass * omega) * np.sin(omega * t))) < 1e-4


def test_reversibility_and_norm(orc):
    psi, v, prm = W.wavepacket_harmonic(seed=11, n=32, step=0.8)
    pot = v.astype(complex)
    step, mass, dt = prm["step"], prm["mass"], prm["dt"]
    fwd = orc.wf_propagate(psi, pot, step, mass, (0, 0, 0), dt)
    back = orc.wf_propagate(fwd, pot, step, mass, (0, 0, 0), -dt)
    assert rel_l2(back, psi) < 1e-13
    assert abs(orc.integral_of_square(fwd, step) - orc.integral_of_square(psi, step)) < 1e-13


def test_imaginary_time_ground_state(orc):
    """Imaginary-time relaxation in V = 1/2 (0.01) r^2 reaches E0 = 3/2 hbar omega = 0.15."""
    n, step, mass = 32, 0.8, 1.0
    shape = (n, n, n)
    v = W.harmonic(shape, step)
    psi = np.ascontiguousarray(W.normalize(W.gaussian_packet(shape, step, 3.0, (0.5, 0, 0), (0, 0, 0)), step))
    pot = v.astype(complex)
    for _ in range(300):
        orc.wf_propagate_inplace(psi, pot, step, mass, (0, 0, 0), -1j * 0.5)
        psi[...] = orc.wf_normalize(psi, step)[0]
    assert abs(orc.energy_fft(psi, v, step, mass) - 0.15) < 2e-4


def test_gpe_uniform_condensate(orc):
    """|psi|^2 = rho0 everywhere: the g|psi|^2 term is the constant g rho0 (phase exp(-i g rho0 t));
    E_GP = hbar^2 k^2 / 2m + (g/2) rho0^2 Vol for a plane wave of unit norm."""
    n, step, mass, g = 16, 0.5, 1.0, 0.7
    shape = (n, n, n)
    vol = (n * step) ** 3
    rho0 = 1.0 / vol
    i, j, k = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    kv = 2 * np.pi / (n * step)
    psi = np.sqrt(rho0) * np.exp(2j * np.pi * (2 * k) / n)  # k_z = 2 kv
    psi = np.ascontiguousarray(psi.astype(np.complex128))
    vext = np.zeros(shape)
    e = orc.gpe_energy(psi, vext, g, step, mass)
    assert abs(e - ((2 * kv) ** 2 / (2 * mass) + 0.5 * g * rho0)) < 1e-12
    b = psi.copy()
    dt = 0.1
    orc.gpe_step_inplace(b, vext, g, step, mass, (0, 0, 0), dt, normalize=False)
    expect = psi * np.exp(-1j * dt * ((2 * kv) ** 2 / (2 * mass) + g * rho0))
    assert rel_l2(b, expect) < 1e-13


def test_gpe_g0_equals_plain_step(orc):
    psi, v, prm = W.wavepacket_harmonic(seed=2, n=32, step=0.8)
    a = psi.copy()
    orc.gpe_step_inplace(a, v, 0.0, prm["step"], 1.0, (0, 0, 0), -0.05j, normalize=True)
    b, _ = orc.wf_normalize(orc.wf_propagate(psi, v.astype(complex), prm["step"], 1.0, (0, 0, 0), -0.05j), prm["step"])
    assert rel_l2(a, b) < 1e-15


def _gauss_overlap_1d(xa, wa, pa, xb, wb, pb):
    """int conj(exp(-(x-xa)^2/wa^2 + i pa (x-xa))) exp(-(x-xb)^2/wb^2 + i pb (x-xb)) dx in closed
    form: exp(-al x^2 + be x + ga) integrates to sqrt(pi/al) exp(be^2/(4 al) + ga)."""
    al = 1 / wa**2 + 1 / wb**2
    be = 2 * xa / wa**2 + 2 * xb / wb**2 + 1j * (pb - pa)
    ga = -xa**2 / wa**2 - xb**2 / wb**2 + 1j * (pa * xa - pb * xb)
    return np.sqrt(np.pi / al) * np.exp(be * be / (4 * al) + ga)


def test_conjugate_product_closed_form_and_vdot(orc):
    """or_integral_of_conjugate_product (src/cgrid.c:1640-1658) is step^d sum conj(a) b: against
    np.vdot on random 3-D, 2-D and 1-D grids, and against the closed-form overlap of two Gaussian
    packets with different centres, widths and momenta (its imaginary part changes sign if the
    conjugate is taken of the second factor, and the step^d factor shows in the modulus)."""
    rng = np.random.default_rng(41)
    for shape, step in (((6, 10, 12), 0.3), ((1, 16, 9), 0.25), ((1, 1, 33), 0.5)):
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        d = sum(1 for n in shape if n > 1)
        ref = np.vdot(a, b) * step**d
        got = orc.integral_of_conjugate_product(a, b, step)
        assert abs(got - ref) < 1e-13 * abs(ref), shape
        assert abs(orc.integral_of_conjugate_product(b, a, step) - np.conj(ref)) < 1e-13 * abs(ref)
    n, step = 96, 0.2
    ra, wa, pa = np.array([0.5, -0.8, 0.3]), 1.3, np.array([0.7, -0.2, 0.4])
    rb, wb, pb = np.array([-0.4, 0.2, 0.9]), 1.6, np.array([-0.3, 0.5, 0.1])
    psi_a = W.gaussian_packet((n, n, n), step, wa, ra, pa)
    psi_b = W.gaussian_packet((n, n, n), step, wb, rb, pb)
    exact = np.prod([_gauss_overlap_1d(ra[m], wa, pa[m], rb[m], wb, pb[m]) for m in range(3)])
    got = orc.integral_of_conjugate_product(psi_a, psi_b, step)
    assert abs(exact.imag) > 0.1 * abs(exact)  # the sign of the imaginary part is tested
    assert abs(got - exact) < 1e-12 * abs(exact)


def _fd_eigen_state(shape, kind, modes, phase=0.3):
    """Exact eigenvectors of the 3-point second difference under each boundary rule, and their
    eigenvalue factor sum_axes (2 cos theta - 2): periodic exp(2 pi i a m / n); Neumann (mirror
    about 0 and n-1) cos(pi a m / (n-1)); zero Dirichlet sin(pi a (m+1) / (n+1))."""
    f = np.ones(shape, dtype=np.complex128) * np.exp(1j * phase)
    lam = 0.0
    for axis, (n, a) in enumerate(zip(shape, modes)):
        m = np.arange(n)
        if kind == "periodic":
            th = 2 * np.pi * a / n
            v = np.exp(1j * th * m)
        elif kind == "neumann":
            th = np.pi * a / (n - 1)
            v = np.cos(th * m)
        else:
            th = np.pi * a / (n + 1)
            v = np.sin(th * (m + 1))
        lam += 2 * np.cos(th) - 2
        sh = [1, 1, 1]
        sh[axis] = n
        f = f * v.reshape(sh)
    return f, lam


@pytest.mark.parametrize("kind,bc", [("periodic", 2), ("neumann", 1), ("dirichlet", 0)])
def test_fd_kinetic_energy_boundary_eigenstates(orc, kind, bc):
    """The FD energy under each boundary rule of src/cgrid_bc.c:21-86 on an exact eigenvector of
    the 7-point Laplacian with that rule: T = -(hbar^2 / 2m) (lambda / h^2) int |psi|^2.  A mirror
    about -1/2 instead of 0 (Neumann) or a wrap for Dirichlet would break the eigenvector."""
    shape, step, mass = (9, 12, 7), 0.35, 1.4
    f, lam = _fd_eigen_state(shape, kind, (2, 3, 1))
    exact = -1.0 / (2 * mass) * lam / step**2 * np.sum(np.abs(f) ** 2) * step**3
    got = orc.kinetic_energy_fd(f, step, mass, bc)
    assert abs(got - exact) < 1e-12 * abs(exact), (kind, got, exact)
    # the periodic wrapper is the periodic rule
    if bc == 2:
        assert orc.kinetic_energy_fd_periodic(f, step, mass) == got


def test_fd_kinetic_energy_dirichlet_value(orc):
    """A non-zero Dirichlet value d (src/cgrid_bc.c:21-24): d + s with s a zero-Dirichlet
    eigenvector has Laplacian lambda s, so T = -(hbar^2/2m)(lambda/h^2)(Re(conj(d) sum s) +
    sum |s|^2) h^3; a constant d has T = 0."""
    shape, step, mass, d = (8, 6, 10), 0.3, 1.0, 0.4 - 0.7j
    s, lam = _fd_eigen_state(shape, "dirichlet", (1, 2, 3))
    f = s + d
    exact = -1.0 / (2 * mass) * lam / step**2 * (np.real(np.conj(d) * s.sum()) + np.sum(np.abs(s) ** 2)) * step**3
    assert abs(orc.kinetic_energy_fd(f, step, mass, 0, d) - exact) < 1e-12 * abs(exact)
    const = np.full(shape, d)
    assert abs(orc.kinetic_energy_fd(const, step, mass, 0, d)) < 1e-13
    # with the default value 0 the constant grid has a boundary jump: not zero
    assert abs(orc.kinetic_energy_fd(const, step, mass, 0, 0.0)) > 1.0


def test_interpolation_oracle_pins(orc):
    """The trilinear interpolation of cgrid_value / rgrid_value (src/cgrid.c:2060-2104,
    src/rgrid.c:2208-2256) against scipy's RegularGridInterpolator inside the grid, exact
    reproduction of a linear function, the grid values at grid points (including negative integer
    coordinates, where the reference's floor takes the lower cell with fraction 1) and the
    periodic wrap between the last and the first point."""
    from scipy.interpolate import RegularGridInterpolator

    rng = np.random.default_rng(5)
    shape, step, origin = (7, 8, 6), 0.5, (0.2, -0.1, 0.3)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    r = rng.standard_normal(shape)
    axes = [orc.axis_coords(n, step, o) for n, o in zip(shape, origin)]
    pts = np.stack([rng.uniform(ax[0] + 1e-9, ax[-1] - 1e-9, 200) for ax in axes], axis=1)
    for vals in (a, r):
        got = orc.interpolate(vals, step, origin, pts[:, 0], pts[:, 1], pts[:, 2], orc.BC_PERIODIC)
        if np.iscomplexobj(vals):
            want = RegularGridInterpolator(axes, vals.real)(pts) + 1j * RegularGridInterpolator(axes, vals.imag)(pts)
        else:
            want = RegularGridInterpolator(axes, vals)(pts)
        assert np.max(np.abs(got - want)) < 1e-13
    x, y, z = np.meshgrid(*axes, indexing="ij")
    lin = 0.3 + 0.7 * x - 1.1 * y + 0.4 * z
    got = orc.interpolate(lin, step, origin, pts[:, 0], pts[:, 1], pts[:, 2], orc.BC_PERIODIC)
    assert np.max(np.abs(got - (0.3 + 0.7 * pts[:, 0] - 1.1 * pts[:, 1] + 0.4 * pts[:, 2]))) < 1e-13
    at_nodes = orc.interpolate(a, step, origin, x, y, z, orc.BC_PERIODIC)
    assert np.max(np.abs(at_nodes - a)) < 1e-14
    # periodic wrap: halfway past the last x point lies between x = nx-1 and x = 0
    xw = axes[0][-1] + 0.5 * step
    got = orc.interpolate(r, step, origin, np.array([xw]), np.array([axes[1][2]]), np.array([axes[2][3]]), orc.BC_PERIODIC)
    assert abs(got[0] - 0.5 * (r[-1, 2, 3] + r[0, 2, 3])) < 1e-14
    # Neumann: halfway before the first x point lies between x = 1 (mirror of -1) and x = 0
    xn = axes[0][0] - 0.5 * step
    got = orc.interpolate(r, step, origin, np.array([xn]), np.array([axes[1][2]]), np.array([axes[2][3]]), orc.BC_NEUMANN)
    assert abs(got[0] - 0.5 * (r[1, 2, 3] + r[0, 2, 3])) < 1e-14


@pytest.mark.parametrize("ns,shift", [(1, 0.0), (2, 0.75), (3, 8.0 / 15.0)])
def test_smooth_map_oracle_closed_form(orc, ns, shift):
    """cgrid_smooth_map's cell average (src/cprivate.h:35-68): ns = 1 is the point value, a linear
    function averages to its centre value (the hat weights are symmetric), and r^2 averages to
    r^2 + 3 <s^2>, where <s^2> = h^2/4 for ns = 2 and 8 h^2/45 for ns = 3."""
    shape, step, origin = (5, 6, 4), 0.4, (0.1, 0.0, -0.2)
    x, y, z = np.meshgrid(*[orc.axis_coords(n, step, o) for n, o in zip(shape, origin)], indexing="ij")

    def quad(x, y, z):
        return x * x + y * y + z * z + 0j

    def lin(x, y, z):
        return 0.5 + 0.2 * x - 0.3 * y + 0.9 * z + 1j * (0.1 * x + 0.4 * z)

    assert np.max(np.abs(orc.smooth_map(lin, shape, step, origin, ns) - lin(x, y, z))) < 1e-14
    got = orc.smooth_map(quad, shape, step, origin, ns)
    assert np.max(np.abs(got - (x * x + y * y + z * z + shift * step * step))) < 1e-14
