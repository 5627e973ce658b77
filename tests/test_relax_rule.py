"""The imaginary-time convergence rule (libgrid_amd/csrc/host/relax_rule.h), compiled with gcc
into a small harness and fed synthetic energy sequences: geometric tails of known limit, with
and without FP32-sized rounding noise.  The rule must stop only once the true remaining error
E_n - E_inf is of the order of tol, including for slow (nearly degenerate) decays."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(HERE, "..", "libgrid_amd", "csrc", "host")

SRC = r"""
#include "relax_rule.h"
/* the step at which the rule stops, -1 if it never does, -2 - k if it reports divergence at k */
static long feed_all(lg_relax_state *s, const double *e, long n, double tol) {
  for (long k = 0; k < n; ++k) {
    const int v = lg_relax_feed(s, k, e[k], tol);
    if (v < 0) return -2 - k;
    if (v) return k;
  }
  return -1;
}
long run_floor(const double *e, long n, double tol, int rise_exit, double lnq_floor) {
  lg_relax_state s;
  lg_relax_init(&s);
  s.rise_exit = rise_exit;
  s.lnq_floor = lnq_floor;
  return feed_all(&s, e, n, tol);
}
long run_ex(const double *e, long n, double tol, int rise_exit) {
  lg_relax_state s;
  lg_relax_init(&s);
  s.rise_exit = rise_exit;
  return feed_all(&s, e, n, tol);
}
long run(const double *e, long n, double tol) { return run_ex(e, n, tol, 1); }
"""


@pytest.fixture(scope="module")
def rule(tmp_path_factory):
    d = tmp_path_factory.mktemp("relax")
    c = d / "r.c"
    c.write_text(SRC)
    so = d / "r.so"
    subprocess.check_call(["gcc", "-O2", "-shared", "-fPIC", "-I", HDR, str(c), "-o", str(so), "-lm"])
    lib = C.CDLL(str(so))
    lib.run_ex.restype = C.c_long
    lib.run_ex.argtypes = [C.POINTER(C.c_double), C.c_long, C.c_double, C.c_int]

    lib.run_floor.restype = C.c_long
    lib.run_floor.argtypes = [C.POINTER(C.c_double), C.c_long, C.c_double, C.c_int, C.c_double]

    def stop(e, tol, rise_exit=1, lnq_floor=0.0):
        a = np.ascontiguousarray(e, dtype=np.float64)
        return lib.run_floor(a.ctypes.data_as(C.POINTER(C.c_double)), len(a), tol, rise_exit, lnq_floor)

    return stop


def tail(rho, n, r0=1e-2, einf=-0.3):
    k = np.arange(n)
    return einf + r0 * rho ** k, einf


@pytest.mark.parametrize("rho", [0.5, 0.9, 0.97, 0.995, 0.9996])
def test_geometric_exact(rule, rho):
    tol = 1e-10
    e, einf = tail(rho, 400000)
    k = rule(e, tol)
    assert k > 0
    rem = e[k] - einf
    # stops with the remaining error below tol, and not needlessly late (within 4x of tol or
    # within one stride of the first step where it holds)
    assert rem < tol
    first = int(np.argmax(e - einf < tol))
    assert rem > tol / 64 or k - first <= max(64, int(2.0 / (1 - rho)))


@pytest.mark.parametrize("rho", [0.97, 0.995, 0.9996])
def test_geometric_fp32_noise(rule, rho):
    """Energies with 5e-9 Gaussian scatter per step (what FP32 states give): a one-step test
    stops orders of magnitude early on slow tails; the windowed rule stays within a few tol."""
    tol = 1e-8
    rng = np.random.default_rng(1)
    clean, einf = tail(rho, 200000)
    worst = 0.0
    for trial in range(20):
        k = rule(clean + rng.normal(0.0, 5e-9, clean.shape), tol)
        assert k > 0
        worst = max(worst, clean[k] - einf)  # true remaining error where it stopped
    print(f"rho {rho}: worst remaining error {worst:.2e}")
    assert worst < 5 * tol


def test_never_stops_with_zero_tol(rule):
    e, _ = tail(0.9, 5000)
    assert rule(e, 0.0) == -1


def test_rise_below_tol_ends(rule):
    # a flat sequence that then rises by less than tol: rounding floor
    e = np.full(100, -0.5)
    e[40:] += 1e-12
    assert 0 < rule(e, 1e-10) < 100


def test_fast_transient_on_slow_tail(rule):
    """E_n - E_inf = a 0.3^n + b 0.999^n with b = 3e-9 >> tol: the fast part hides the slow tail
    from a single check (q ~ 0.3^m); the rule must still relax the slow part below a few tol."""
    tol = 1e-10
    k = np.arange(300000)
    for a in (1e-6, 1e-8, 3e-9):
        e = -0.2 + a * 0.3 ** k + 3e-9 * 0.999 ** k
        n = rule(e, tol)
        assert n > 0
        rem = e[n] + 0.2
        print(f"a {a:.0e}: stopped at {n}, remaining {rem:.2e}")
        assert rem < 3 * tol


@pytest.mark.parametrize("rho", [0.97, 0.995, 0.9996])
def test_tail_from_below_after_extrapolation(rule, rho):
    """An extrapolation that overshoots a slow mode leaves E_n = E_inf - a rho^n, a rise.  With
    rise_exit = 0 the rule treats it as a geometric tail and stops only once |E_n - E_inf| is
    below tol; the default rule (rise_exit = 1) would take the first small rise for rounding
    and stop with the error still 3e-10."""
    tol = 1e-10
    k = np.arange(400000)
    einf = -0.08
    e = einf - 3e-10 * rho ** k
    assert rule(e, tol, 1) == 32  # the default: stops at the third checkpoint
    s = rule(e, tol, 0)
    assert s > 0 and abs(e[s] - einf) < tol
    # the same for a tail from above, unchanged by the flag
    e2 = einf + 3e-10 * rho ** k
    s2 = rule(e2, tol, 0)
    assert s2 == rule(e2, tol, 1) and e2[s2] - einf < tol


def test_rate_floor_after_extrapolation(rule):
    """After an extrapolation a fast component it excited (1e-7, rho 0.95) sits on what is left
    of the slow mode (3e-10, rho 0.9994).  Without a floor the ratios of 16-step windows are the
    fast one's and the rule stops with 3e-10 still to come; with the slow rate as the floor it
    waits until the remainder is below tol."""
    tol = 1e-10
    k = np.arange(200000)
    einf = -0.08
    e = einf + 1e-7 * 0.95 ** k + 3e-10 * 0.9994 ** k
    early = rule(e, tol, 0)
    assert e[early] - einf > 2e-10
    s = rule(e, tol, 0, np.log(0.9994))
    assert e[s] - einf < tol


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), float("-inf")])
@pytest.mark.parametrize("at", [0, 5, 16, 40])
def test_non_finite_energy_is_divergence(rule, bad, at):
    """A non-finite energy ends the relaxation as diverged (-1) at the step it appears, whether or
    not that step is a checkpoint; before it the rule behaves as usual."""
    e = 1.0 + 0.5 ** np.arange(64.0)
    e[at] = bad
    assert rule(e, 1e-12) == -2 - at
