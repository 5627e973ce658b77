/*
 * Convergence rule of imaginary-time relaxation, shared by the FP64 relaxation on the host
 * (grid_wf.c) and the batched FP32 kernel (hip/batch2d.hpp, which defines LG_RELAX_FN as a
 * device function qualifier before including this file).
 *
 * Imaginary time converges geometrically, E_n - E_inf ~ R rho^n.  The energies are sampled at
 * checkpoints of stride m; with the decreases D1 = E(n-2m) - E(n-m) and D2 = E(n-m) - E(n) the
 * ratio q = D2 / D1 estimates rho^m and E(n) - E_inf ~ D2 q / (1 - q).  The relaxation stops
 * when D2 < tol and that remaining error is below tol.  While q > 1/2 the stride doubles (up
 * to LG_RELAX_MAX_STRIDE), so a window always spans about one decay time and its decrease is
 * of the order of the remaining error rather than of the rounding noise of a single step (FP32
 * energies scatter by a few 1e-9 per step, which makes one-step ratios meaningless near
 * 1e-8).  A decrease D2 that is negative and smaller than tol means rounding dominates the
 * energy change: that also ends the relaxation.
 *
 * The estimate is trusted only when two successive ratios of one stride agree within a factor
 * of two (the larger one is used), and windows are at least LG_RELAX_MIN_STRIDE steps long.
 * Both guard against a fast transient ending on top of a slow tail (E_n - E_inf = a rho_1^n +
 * b rho_2^n, rho_1 << rho_2): short windows see q ~ rho_1^m while b is still far above tol,
 * and the slow part's decrease per window can be below tol.  An FP64 stage started from an
 * FP32-relaxed state looks exactly like that.
 *
 * With rise_exit = 0 (set by the caller after it has extrapolated the state in FP64, where
 * rounding is far below tol) a tail that approaches from below, d1 < d2 < 0, is treated as
 * geometric like one from above instead of ending the relaxation as a small rise: the energy
 * is first-order in the state's error, and an extrapolation that overshoots a mode changes the
 * sign of that error.
 *
 * The caller may also give a floor for the ratio, as a decay rate per step (lnq_floor < 0):
 * after an extrapolation the slowest mode is known to decay by r per s steps, so what is left
 * of it decays no faster, and q = exp(lnq_floor m) bounds the ratio of a window of m steps
 * from below.  Without the floor a fast component that the extrapolation excited could set
 * q while the slow remainder's decrease per window is still below tol.
 */
#ifndef LG_RELAX_RULE_H
#define LG_RELAX_RULE_H

#include <math.h>

#ifndef LG_RELAX_FN
#define LG_RELAX_FN static inline
#endif

#define LG_RELAX_MIN_STRIDE 16
#define LG_RELAX_MAX_STRIDE 4096

typedef struct {
  double c0, c1; /* energies at the two previous checkpoints */
  double qprev;  /* ratio at the previous check of this stride (1 if none) */
  long m;        /* stride */
  long next;     /* step of the next checkpoint */
  int nchk;      /* checkpoints taken at this stride (0, 1, 2) */
  int rise_exit; /* 1: a rise smaller than tol ends the relaxation (rounding); 0: see above */
  double lnq_floor; /* 0: no floor; < 0: ln of the slowest known decay per step */
} lg_relax_state;

LG_RELAX_FN void lg_relax_init(lg_relax_state *s) {
  s->c0 = s->c1 = 0.0;
  s->qprev = 1.0;
  s->m = LG_RELAX_MIN_STRIDE;
  s->next = 0;
  s->nchk = 0;
  s->rise_exit = 1;
  s->lnq_floor = 0.0;
}

/* Feed E_n, the energy of the state entering step n (n = 0, 1, 2, ...); returns 1 once the
 * energy is converged to tol, and -1 at once if E_n is not finite (the relaxation diverged: the
 * counterpart of the reference's floating-point exception trap, src/threads.c:122-126). */
LG_RELAX_FN int lg_relax_feed(lg_relax_state *s, long n, double e, double tol) {
  if (!isfinite(e)) return -1;
  if (n != s->next) return 0;
  s->next = n + s->m;
  if (s->nchk < 2) {
    if (s->nchk == 0) s->c0 = e;
    else s->c1 = e;
    ++s->nchk;
    return 0;
  }
  const double d1 = s->c0 - s->c1, d2 = s->c1 - e;
  const int up = !s->rise_exit && d1 < 0.0 && d2 < 0.0 && d2 > d1;
  const int geometric = (d1 > 0.0 && d2 > 0.0 && d2 < d1) || up;
  const double q = geometric ? d2 / d1 : 1.0;
  const double a2 = d2 < 0.0 ? -d2 : d2;
  /* a small rise ends it as rounding; with rise_exit = 0 only right after a decrease */
  if (d2 <= 0.0 && -d2 < tol && (s->rise_exit || d1 > 0.0)) return 1;
  const int consistent = geometric && s->qprev < 1.0 && q < 2.0 * s->qprev && s->qprev < 2.0 * q;
  double qq = q > s->qprev ? q : s->qprev;
  if (s->lnq_floor < 0.0) {
    const double qf = exp(s->lnq_floor * (double)s->m);
    if (qf > qq) qq = qf;
  }
  if (consistent && a2 < tol && a2 * qq < tol * (1.0 - qq)) return 1;
  s->qprev = q;
  if (((d1 > 0.0 && d2 > 0.0) || up) && q > 0.5 && s->m < LG_RELAX_MAX_STRIDE) {
    /* double the stride: the checkpoints E(n - 2m) and E(n) are the new pair */
    s->c1 = e;
    s->m *= 2;
    s->next = n + s->m;
    s->qprev = 1.0;
    return 0;
  }
  s->c0 = s->c1;
  s->c1 = e;
  return 0;
}

#endif
