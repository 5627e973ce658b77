/*
 * ORACLE — test infrastructure only.  A CPU (C99 + OpenMP) restatement of the reference
 * library's hot path, used by tests/ (as the checker), by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  Nothing in libgrid_amd/ links or calls it.
 *
 * Each function states the reference code it follows.  The reference transforms with FFTW
 * (src/cgrid_fft.c:22-146); FFTW is not available here, so the FFT below is this oracle's
 * own mixed-radix Cooley-Tukey with the same conventions: forward = exp(-2 pi i ...),
 * backward = exp(+2 pi i ...), both unnormalised, in place, row-major [nx][ny][nz].
 * Twiddles are evaluated in x87 extended precision and rounded once.
 *
 * Pinning (see tests/test_oracle.py): FFT against numpy.fft golden vectors and brute-force
 * DFTs; kinetic phase against the closed-form free-particle evolution; potential phase
 * against constant / linear / absorbing potentials; norm and energies against analytic
 * Gaussian integrals; the full step against Ehrenfest dynamics, time reversibility and the
 * harmonic-oscillator ground state; the GPE composition against the same invariants with
 * g = 0 and against a first-order expansion in g; the cut-off (CFFT) kinetic phase against its
 * limits, the Nyquist convention of its test and plane-wave eigenstates; the 4th-order step
 * against the 2nd-order one without a potential, constant potentials, reversibility and the
 * closed-form imaginary-time fixed point of its stage sequence in a harmonic trap; the
 * absorbing boundary against hand-computed amplitudes, its alpha = 0 / alpha = 1 regions and
 * the absorption of an outgoing packet.
 *
 * Reductions: the reference sums with OpenMP reduction(+) in thread order; the oracle sums
 * each [nx][ny] row in index order and the row sums with Neumaier compensation, so results
 * do not depend on the thread count.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

#ifndef HBAR
#define HBAR 1.0
#endif

/* ---- compensated accumulation ------------------------------------------------------------ */

typedef struct {
  double s, c;
} ksum;
static inline void kadd(ksum *k, double x) {
  double t = k->s + x;
  if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
  else k->c += (x - t) + k->s;
  k->s = t;
}

/* ---- FFT ----------------------------------------------------------------------------------- */

typedef struct {
  long n;
  int sign;
  cplx *tw; /* exp(sign 2 pi i j / n), j < n */
} fft_plan;

static void make_twiddles(fft_plan *p, long n, int sign) {
  p->n = n;
  p->sign = sign;
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (long j = 0; j < n; ++j) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
    double re = (double)cosl(a), im = (double)sinl(a);
    if (j == 0) { re = 1.0; im = 0.0; }
    else if (4 * j == n) { re = 0.0; im = 1.0; }
    else if (2 * j == n) { re = -1.0; im = 0.0; }
    else if (4 * j == 3 * n) { re = 0.0; im = -1.0; }
    p->tw[j] = re + (double)sign * im * I;
  }
}

static long smallest_factor(long n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (long f = 3; f * f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

/* out[0..n) = DFT of in[0], in[s], ..., in[(n-1)s]; tw indexes the full-size table with
 * stride ts (W_n^e = tw[e * ts]). */
static void fft_rec(const cplx *in, cplx *out, long n, long s, const cplx *tw, long ts, long N) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  long p = smallest_factor(n), m = n / p;
  for (long r = 0; r < p; ++r) fft_rec(in + r * s, out + r * m, m, s * p, tw, ts * p, N);
  cplx t[64], *tbig = NULL, *tt = t;
  if (p > 64) tt = tbig = (cplx *)malloc(sizeof(cplx) * (size_t)p);
  for (long k = 0; k < m; ++k) {
    for (long r = 0; r < p; ++r) tt[r] = r ? out[r * m + k] * tw[(r * k * ts) % N] : out[k];
    if (p == 2) {
      out[k] = tt[0] + tt[1];
      out[k + m] = tt[0] - tt[1];
    } else if (p == 4) {
      /* W_4 = tw[N/4] (= -i forward, +i backward) */
      cplx w = tw[N / 4];
      cplx a0 = tt[0] + tt[2], a1 = tt[0] - tt[2], a2 = tt[1] + tt[3], a3 = (tt[1] - tt[3]) * w;
      out[k] = a0 + a2;
      out[k + m] = a1 + a3;
      out[k + 2 * m] = a0 - a2;
      out[k + 3 * m] = a1 - a3;
    } else {
      for (long q = 0; q < p; ++q) {
        cplx acc = 0;
        for (long r = 0; r < p; ++r) acc += tt[r] * tw[((r * q) % p) * (N / p)];
        out[k + q * m] = acc;
      }
    }
  }
  free(tbig);
}

static void fft1d(fft_plan *pl, cplx *data, long stride, cplx *buf_in, cplx *buf_out) {
  long n = pl->n;
  for (long i = 0; i < n; ++i) buf_in[i] = data[i * stride];
  fft_rec(buf_in, buf_out, n, 1, pl->tw, 1, n);
  for (long i = 0; i < n; ++i) data[i * stride] = buf_out[i];
}

/* In-place unnormalised 3-D DFT, sign -1 forward / +1 backward (reference src/cgrid_fft.c:48-52,
 * fftw_plan_dft_3d(nx, ny, nz, ..., FFTW_FORWARD / FFTW_BACKWARD)). */
void or_fft3d(cplx *a, long nx, long ny, long nz, int sign) {
  long dims[3] = {nx, ny, nz};
  for (int axis = 2; axis >= 0; --axis) {
    long n = dims[axis];
    if (n == 1) continue;
    fft_plan pl;
    make_twiddles(&pl, n, sign);
    long stride = axis == 2 ? 1 : axis == 1 ? nz : ny * nz;
    long outer = axis == 0 ? 1 : nx, inner = axis == 2 ? ny : nz, ostride = axis == 0 ? 0 : ny * nz;
    long istride = axis == 2 ? nz : 1;
    if (axis == 0) { outer = 1; inner = ny * nz; istride = 1; }
    long npen = outer * inner;
#pragma omp parallel
    {
      cplx *b1 = (cplx *)malloc(sizeof(cplx) * (size_t)n), *b2 = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
      for (long q = 0; q < npen; ++q) {
        long o = q / inner, i = q % inner;
        fft1d(&pl, a + o * ostride + i * istride, stride, b1, b2);
      }
      free(b1);
      free(b2);
    }
    free(pl.tw);
  }
}

/* ---- cgrid wrappers (reference src/cgrid.c:1822-1890) ------------------------------------- */

void or_cgrid_fft(cplx *a, long nx, long ny, long nz) { or_fft3d(a, nx, ny, nz, -1); }
void or_cgrid_inverse_fft(cplx *a, long nx, long ny, long nz) { or_fft3d(a, nx, ny, nz, +1); }

/* cgrid_multiply (src/cgrid.c:1223-1238) */
void or_cgrid_multiply(cplx *a, long n, double cre, double cim) {
  cplx c = cre + cim * I;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) a[i] *= c;
}

/* cgrid_inverse_fft_norm = inverse FFT then multiply by fft_norm = 1/(nx ny nz) (:1857-1876, :85) */
void or_cgrid_inverse_fft_norm(cplx *a, long nx, long ny, long nz) {
  or_fft3d(a, nx, ny, nz, +1);
  or_cgrid_multiply(a, nx * ny * nz, 1.0 / (double)(nx * ny * nz), 0.0);
}

static double step_factor(long nx, long ny, long nz, double step) {
  double f = 1.0;
  if (nx != 1) f *= step;
  if (ny != 1) f *= step;
  if (nz != 1) f *= step;
  return f;
}

/* cgrid_integral_of_square (src/cgrid.c:1610-1629): step^d * sum |a|^2 */
double or_integral_of_square(const cplx *a, long nx, long ny, long nz, double step) {
  long nxy = nx * ny;
  double *rows = (double *)malloc(sizeof(double) * (size_t)nxy);
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    double s = 0;
    for (long k = 0; k < nz; ++k) {
      cplx v = a[ij * nz + k];
      s += creal(v) * creal(v) + cimag(v) * cimag(v);
    }
    rows[ij] = s;
  }
  ksum k = {0, 0};
  for (long ij = 0; ij < nxy; ++ij) kadd(&k, rows[ij]);
  free(rows);
  return (k.s + k.c) * step_factor(nx, ny, nz, step);
}

/* cgrid_integral_of_conjugate_product (src/cgrid.c:1640-1658) */
void or_integral_of_conjugate_product(const cplx *a, const cplx *b, long nx, long ny, long nz, double step, double *out2) {
  long nxy = nx * ny;
  double *rr = (double *)malloc(sizeof(double) * (size_t)nxy), *ri = (double *)malloc(sizeof(double) * (size_t)nxy);
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    cplx s = 0;
    for (long k = 0; k < nz; ++k) s += conj(a[ij * nz + k]) * b[ij * nz + k];
    rr[ij] = creal(s);
    ri[ij] = cimag(s);
  }
  ksum kr = {0, 0}, ki = {0, 0};
  for (long ij = 0; ij < nxy; ++ij) {
    kadd(&kr, rr[ij]);
    kadd(&ki, ri[ij]);
  }
  free(rr);
  free(ri);
  double f = step_factor(nx, ny, nz, step);
  out2[0] = (kr.s + kr.c) * f;
  out2[1] = (ki.s + ki.c) * f;
}

/* cgrid_fft_laplace_expectation_value (src/cgrid_diff.c:817-879): -sum k^2 |a_k|^2 * fft_norm2,
 * k_x = 2 pi i / (nx step) for i <= nx/2 else 2 pi (i - nx) / (nx step), no k0 shift. */
double or_fft_laplace_expectation_value(const cplx *a, long nx, long ny, long nz, double step) {
  long nxy = nx * ny, nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
  double lx = 2.0 * M_PI / (step * (double)nx), ly = 2.0 * M_PI / (step * (double)ny), lz = 2.0 * M_PI / (step * (double)nz);
  double norm = step_factor(nx, ny, nz, step) / (double)(nx * ny * nz);
  double *rows = (double *)malloc(sizeof(double) * (size_t)nxy);
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    long i = ij / ny, j = ij % ny;
    double kx = (i <= nx2) ? (double)i * lx : (double)(i - nx) * lx;
    double ky = (j <= ny2) ? (double)j * ly : (double)(j - ny) * ly;
    double ssum = 0.0;
    for (long k = 0; k < nz; ++k) {
      double kz = (k <= nz2) ? (double)k * lz : (double)(k - nz) * lz;
      cplx v = a[ij * nz + k];
      ssum -= (kx * kx + ky * ky + kz * kz) * (creal(v) * creal(v) + cimag(v) * cimag(v));
    }
    rows[ij] = ssum;
  }
  ksum s = {0, 0};
  for (long ij = 0; ij < nxy; ++ij) kadd(&s, rows[ij]);
  free(rows);
  return (s.s + s.c) * norm;
}

/* grid_grid_expectation_value (src/ngrid.c:355-376): step^d sum V |a|^2, V dense [nx][ny][nz] */
double or_grid_expectation_value(const cplx *a, const double *v, long nx, long ny, long nz, double step) {
  long nxy = nx * ny;
  double *rows = (double *)malloc(sizeof(double) * (size_t)nxy);
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    double s = 0;
    for (long k = 0; k < nz; ++k) {
      cplx x = a[ij * nz + k];
      s += (creal(x) * creal(x) + cimag(x) * cimag(x)) * v[ij * nz + k];
    }
    rows[ij] = s;
  }
  ksum k = {0, 0};
  for (long ij = 0; ij < nxy; ++ij) kadd(&k, rows[ij]);
  free(rows);
  return (k.s + k.c) * step_factor(nx, ny, nz, step);
}

/* ---- wavefunction operators ---------------------------------------------------------------- */

/* grid_wf_propagate_kinetic_fft (src/grid_wf_fft.c:141-193): a is in k-space;
 * a *= cexp(time_mass * (kx^2 + ky^2 + kz^2)), time_mass = -i time hbar / (2 m), k - k0. */
void or_propagate_kinetic_fft(cplx *a, long nx, long ny, long nz, double step, double mass, double kx0, double ky0,
                              double kz0, double time_re, double time_im) {
  long nxy = nx * ny, nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
  double lx = 2.0 * M_PI / (step * (double)nx), ly = 2.0 * M_PI / (step * (double)ny), lz = 2.0 * M_PI / (step * (double)nz);
  cplx time = time_re + time_im * I;
  cplx time_mass = -I * time * HBAR / (mass * 2.0);
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    long i = ij / ny, j = ij % ny;
    double kx = (i <= nx2) ? ((double)i) * lx - kx0 : ((double)(i - nx)) * lx - kx0;
    double ky = (j <= ny2) ? ((double)j) * ly - ky0 : ((double)(j - ny)) * ly - ky0;
    for (long k = 0; k < nz; ++k) {
      double kz = (k <= nz2) ? ((double)k) * lz - kz0 : ((double)(k - nz)) * lz - kz0;
      a[ij * nz + k] *= cexp(time_mass * (kx * kx + ky * ky + kz * kz));
    }
  }
}

/* grid_wf_propagate_kinetic_cfft (src/grid_wf_fft.c:206-262): as the kinetic phase above,
 * but a component with tot = kx^2 + ky^2 + kz^2 >= kmax^2 is set to zero (k - k0, the same
 * Nyquist convention, tot summed in the order x, y, z). */
void or_propagate_kinetic_cfft(cplx *a, long nx, long ny, long nz, double step, double mass, double kx0, double ky0,
                               double kz0, double time_re, double time_im, double kmax) {
  long nxy = nx * ny, nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
  double lx = 2.0 * M_PI / (step * (double)nx), ly = 2.0 * M_PI / (step * (double)ny), lz = 2.0 * M_PI / (step * (double)nz);
  double kmax2 = kmax * kmax;
  cplx time = time_re + time_im * I;
  cplx time_mass = -I * time * HBAR / (mass * 2.0);
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    long i = ij / ny, j = ij % ny;
    double kx = (i <= nx2) ? ((double)i) * lx - kx0 : ((double)(i - nx)) * lx - kx0;
    double ky = (j <= ny2) ? ((double)j) * ly - ky0 : ((double)(j - ny)) * ly - ky0;
    for (long k = 0; k < nz; ++k) {
      double kz = (k <= nz2) ? ((double)k) * lz - kz0 : ((double)(k - nz)) * lz - kz0;
      double tot = kx * kx + ky * ky + kz * kz;
      if (tot < kmax2) a[ij * nz + k] *= cexp(time_mass * tot);
      else a[ij * nz + k] = 0.0;
    }
  }
}

/* grid_wf_propagate_potential (src/grid_wf.c:569-600), FFT propagator (no ts_func):
 * a *= cexp(c * (cons + pot)), c = -i tstep / hbar, pot complex */
void or_propagate_potential(cplx *a, const cplx *pot, long n, double tre, double tim, double cons) {
  cplx c = -I * (tre + tim * I) / HBAR;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) a[i] = a[i] * cexp(c * (cons + pot[i]));
}

/* grid_wf_propagate, WF_2ND_ORDER_FFT, no absorbing boundary (src/grid_wf.c:379-432):
 * fft, kinetic(t/2), inverse_fft_norm, potential(t, cons), fft, kinetic(t/2), inverse_fft_norm */
void or_wf_propagate(cplx *a, const cplx *pot, long nx, long ny, long nz, double step, double mass, double kx0, double ky0,
                     double kz0, double tre, double tim) {
  double cons = -(HBAR * HBAR / (2.0 * mass)) * (kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
  long n = nx * ny * nz;
  or_cgrid_fft(a, nx, ny, nz);
  or_propagate_kinetic_fft(a, nx, ny, nz, step, mass, kx0, ky0, kz0, 0.5 * tre, 0.5 * tim);
  or_cgrid_inverse_fft_norm(a, nx, ny, nz);
  if (pot) or_propagate_potential(a, pot, n, tre, tim, cons);
  or_cgrid_fft(a, nx, ny, nz);
  or_propagate_kinetic_fft(a, nx, ny, nz, step, mass, kx0, ky0, kz0, 0.5 * tre, 0.5 * tim);
  or_cgrid_inverse_fft_norm(a, nx, ny, nz);
}

/* grid_wf_propagate, WF_2ND_ORDER_CFFT, no absorbing boundary (src/grid_wf.c:491-498): the
 * WF_2ND_ORDER_FFT sequence with the cut-off kinetic phase */
void or_wf_propagate_cfft(cplx *a, const cplx *pot, long nx, long ny, long nz, double step, double mass, double kx0,
                          double ky0, double kz0, double tre, double tim, double kmax) {
  double cons = -(HBAR * HBAR / (2.0 * mass)) * (kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
  long n = nx * ny * nz;
  or_cgrid_fft(a, nx, ny, nz);
  or_propagate_kinetic_cfft(a, nx, ny, nz, step, mass, kx0, ky0, kz0, 0.5 * tre, 0.5 * tim, kmax);
  or_cgrid_inverse_fft_norm(a, nx, ny, nz);
  if (pot) or_propagate_potential(a, pot, n, tre, tim, cons);
  or_cgrid_fft(a, nx, ny, nz);
  or_propagate_kinetic_cfft(a, nx, ny, nz, step, mass, kx0, ky0, kz0, 0.5 * tre, 0.5 * tim, kmax);
  or_cgrid_inverse_fft_norm(a, nx, ny, nz);
}

/* grid_wf_propagate, WF_4TH_ORDER_FFT / WF_4TH_ORDER_CFFT without an absorbing boundary
 * (src/grid_wf.c:433-457, :505-525): V(t/6), K(t/2), V(2t/3), K(t/2), V(t/6), each V with the
 * same cons as the 2nd-order step.  The reference also forms (1/48)(hbar^2/m)|t|^2 |grad V|^2
 * in a workspace (:449-451) but passes the unmodified potential to the V(2t/3) stage (:452),
 * so that term never reaches psi and is not formed here.  kmax < 0: plain kinetic phase;
 * kmax >= 0: the cut-off one. */
void or_wf_propagate_4th(cplx *a, const cplx *pot, long nx, long ny, long nz, double step, double mass, double kx0,
                         double ky0, double kz0, double tre, double tim, double kmax) {
  double cons = -(HBAR * HBAR / (2.0 * mass)) * (kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
  long n = nx * ny * nz;
  cplx t = tre + tim * I, t6 = t / 6.0, t23 = 2.0 * t / 3.0, th = 0.5 * t;
  for (int half = 0; half < 2; ++half) {
    if (pot) or_propagate_potential(a, pot, n, creal(half ? t23 : t6), cimag(half ? t23 : t6), cons);
    or_cgrid_fft(a, nx, ny, nz);
    if (kmax < 0.0) or_propagate_kinetic_fft(a, nx, ny, nz, step, mass, kx0, ky0, kz0, creal(th), cimag(th));
    else or_propagate_kinetic_cfft(a, nx, ny, nz, step, mass, kx0, ky0, kz0, creal(th), cimag(th), kmax);
    or_cgrid_inverse_fft_norm(a, nx, ny, nz);
  }
  if (pot) or_propagate_potential(a, pot, n, creal(t6), cimag(t6), cons);
}

/* grid_wf_absorb (src/grid_wf.c:185-203): the absorbing-boundary amplitude, 0 inside
 * [lx, hx] x [ly, hy] x [lz, hz] and rising linearly outside; both sides of an axis are
 * divided by the lower bound (lx, ly, lz), as the reference does; the sum over the three
 * axes is multiplied by 2/3 and capped at 1. */
double or_wf_absorb(long i, long j, long k, long lx, long hx, long ly, long hy, long lz, long hz) {
  double t = 0.0;
  if (i < lx) t += ((double)(lx - i)) / (double)lx;
  else if (i > hx) t += ((double)(i - hx)) / (double)lx;
  if (j < ly) t += ((double)(ly - j)) / (double)ly;
  else if (j > hy) t += ((double)(j - hy)) / (double)ly;
  if (k < lz) t += ((double)(lz - k)) / (double)lz;
  else if (k > hz) t += ((double)(k - hz)) / (double)lz;
  t *= 2.0 / 3.0;
  if (t > 1.0) return 1.0;
  return t;
}

/* grid_wf_merge (src/grid_wf.c:728-755): dst = (1 - alpha) wr + alpha wi, alpha from
 * grid_wf_absorb; dst may be wr */
void or_wf_merge(cplx *dst, const cplx *wr, const cplx *wi, long nx, long ny, long nz, long lx, long hx, long ly, long hy,
                 long lz, long hz) {
  long nxy = nx * ny;
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    long i = ij / ny, j = ij % ny;
    for (long k = 0; k < nz; ++k) {
      double alpha = or_wf_absorb(i, j, k, lx, hx, ly, hy, lz, hz);
      dst[ij * nz + k] = (1.0 - alpha) * wr[ij * nz + k] + alpha * wi[ij * nz + k];
    }
  }
}

/* grid_wf_propagate with an absorbing boundary, WF_2ND_ORDER_FFT / _CFFT and Re(time) != 0
 * (src/grid_wf.c:396-422, :464-490): the state is propagated once with time and once (a copy)
 * with -i Re(time), and the two are merged with grid_wf_merge.  kmax < 0: plain kinetic
 * phase; kmax >= 0: the cut-off one. */
void or_wf_propagate_absorb(cplx *a, const cplx *pot, long nx, long ny, long nz, double step, double mass, double kx0,
                            double ky0, double kz0, double tre, double tim, double kmax, long lx, long hx, long ly,
                            long hy, long lz, long hz) {
  long n = nx * ny * nz;
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  memcpy(w, a, sizeof(cplx) * (size_t)n);
  if (kmax < 0.0) {
    or_wf_propagate(a, pot, nx, ny, nz, step, mass, kx0, ky0, kz0, tre, tim);
    or_wf_propagate(w, pot, nx, ny, nz, step, mass, kx0, ky0, kz0, 0.0, -tre);
  } else {
    or_wf_propagate_cfft(a, pot, nx, ny, nz, step, mass, kx0, ky0, kz0, tre, tim, kmax);
    or_wf_propagate_cfft(w, pot, nx, ny, nz, step, mass, kx0, ky0, kz0, 0.0, -tre, kmax);
  }
  or_wf_merge(a, a, w, nx, ny, nz, lx, hx, ly, hy, lz, hz);
  free(w);
}

/* grid_wf_normalize (src/grid_wf.c:694-700): multiply by sqrt(target / norm); returns norm */
double or_wf_normalize(cplx *a, long nx, long ny, long nz, double step, double target) {
  double norm = or_integral_of_square(a, nx, ny, nz, step);
  or_cgrid_multiply(a, nx * ny * nz, sqrt(target / norm), 0.0);
  return norm;
}

/* grid_wf_kinetic_energy_fft (src/grid_wf_fft.c:119-129): on a copy, fft, then
 * -hbar^2/(2m) * laplace expectation value (the reference transforms back in place) */
double or_kinetic_energy_fft(const cplx *a, long nx, long ny, long nz, double step, double mass) {
  long n = nx * ny * nz;
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  memcpy(w, a, sizeof(cplx) * (size_t)n);
  or_cgrid_fft(w, nx, ny, nz);
  double t = -HBAR * HBAR / (2.0 * mass) * or_fft_laplace_expectation_value(w, nx, ny, nz, step);
  free(w);
  return t;
}

/* grid_wf_energy_fft (src/grid_wf_fft.c:101-109) */
double or_energy_fft(const cplx *a, const double *v, long nx, long ny, long nz, double step, double mass) {
  double e = or_kinetic_energy_fft(a, nx, ny, nz, step, mass);
  if (v) e += or_grid_expectation_value(a, v, nx, ny, nz, step);
  return e;
}

/* The value at (i, j, k) under a grid's boundary function (src/cgrid.c:1960-1965 with
 * src/cgrid_bc.c:21-86): bc 2 periodic (wrap at 0 and n), 1 Neumann (mirror about 0 and n - 1;
 * a mirrored index still outside, possible only for n < 3, falls back to the Dirichlet value),
 * 0 Dirichlet (the constant dir). */
static cplx fd_value(const cplx *a, long nx, long ny, long nz, long i, long j, long k, int bc, cplx dir) {
  if (i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz) return a[(i * ny + j) * nz + k];
  if (bc == 2) {
    i %= nx; if (i < 0) i += nx;
    j %= ny; if (j < 0) j += ny;
    k %= nz; if (k < 0) k += nz;
    return a[(i * ny + j) * nz + k];
  }
  if (bc == 1) {
    if (i < 0) i = -i;
    if (i > nx - 1) i = 2 * nx - i - 2;
    if (j < 0) j = -j;
    if (j > ny - 1) j = 2 * ny - j - 2;
    if (k < 0) k = -k;
    if (k > nz - 1) k = 2 * nz - k - 2;
    if (i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz) return a[(i * ny + j) * nz + k];
  }
  return dir;
}

/* finite-difference kinetic energy, grid_wf_kinetic_energy_cn (src/grid_wf_cn.c:37-49):
 * Re int conj(psi) (-hbar^2/2m) lap7(psi), with cgrid_fd_laplace (src/cgrid_diff.c:303-330) and
 * the boundary rule bc of fd_value; dir_re + i dir_im is the Dirichlet value. */
double or_kinetic_energy_fd(const cplx *a, long nx, long ny, long nz, double step, double mass, int bc, double dir_re,
                            double dir_im) {
  long nxy = nx * ny;
  double inv_delta2 = 1.0 / (step * step);
  const cplx dir = dir_re + I * dir_im;
  double *rows = (double *)malloc(sizeof(double) * (size_t)nxy);
#define AT(i, j, k) fd_value(a, nx, ny, nz, (i), (j), (k), bc, dir)
#pragma omp parallel for schedule(static)
  for (long ij = 0; ij < nxy; ++ij) {
    long i = ij / ny, j = ij % ny;
    double s = 0;
    for (long k = 0; k < nz; ++k) {
      cplx lap = inv_delta2 * (-6.0 * AT(i, j, k) + AT(i, j, k + 1) + AT(i, j, k - 1) + AT(i, j + 1, k) + AT(i, j - 1, k) +
                               AT(i + 1, j, k) + AT(i - 1, j, k));
      lap *= -HBAR * HBAR / (2.0 * mass);
      s += creal(conj(AT(i, j, k)) * lap);
    }
    rows[ij] = s;
  }
#undef AT
  ksum k = {0, 0};
  for (long ij = 0; ij < nxy; ++ij) kadd(&k, rows[ij]);
  free(rows);
  return (k.s + k.c) * step_factor(nx, ny, nz, step);
}

double or_kinetic_energy_fd_periodic(const cplx *a, long nx, long ny, long nz, double step, double mass) {
  return or_kinetic_energy_fd(a, nx, ny, nz, step, mass, 2, 0.0, 0.0);
}

/* ---- Gross-Pitaevskii composition of reference calls ---------------------------------------
 * grid_wf_density (src/grid_wf.c:611-627) -> V_eff = vext + g rho -> grid_real_to_complex_re
 * (src/ngrid.c:53-80) -> grid_wf_propagate -> [grid_wf_normalize].  kmax < 0 selects the
 * WF_2ND_ORDER_FFT step, kmax >= 0 the WF_2ND_ORDER_CFFT step with that cutoff. */
void or_gpe_step(cplx *a, const double *vext, double g, long nx, long ny, long nz, double step, double mass, double kx0,
                 double ky0, double kz0, double tre, double tim, int normalize, double target, double kmax) {
  long n = nx * ny * nz;
  cplx *pot = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double rho = creal(conj(a[i]) * a[i]);
    pot[i] = (cplx)(vext[i] + g * rho);
  }
  if (kmax < 0.0) or_wf_propagate(a, pot, nx, ny, nz, step, mass, kx0, ky0, kz0, tre, tim);
  else or_wf_propagate_cfft(a, pot, nx, ny, nz, step, mass, kx0, ky0, kz0, tre, tim, kmax);
  if (normalize) or_wf_normalize(a, nx, ny, nz, step, target);
  free(pot);
}

/* E_GP = T_fft + int vext |a|^2 + (g/2) int |a|^4 */
double or_gpe_energy(const cplx *a, const double *vext, double g, long nx, long ny, long nz, double step, double mass) {
  long n = nx * ny * nz;
  double e = or_energy_fft(a, vext, nx, ny, nz, step, mass);
  double *rho2 = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double r = creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
    rho2[i] = r;
  }
  double q = 0;
  {
    long nxy = nx * ny;
    double *rows = (double *)malloc(sizeof(double) * (size_t)nxy);
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < nxy; ++ij) {
      double s = 0;
      for (long k = 0; k < nz; ++k) s += rho2[ij * nz + k] * rho2[ij * nz + k];
      rows[ij] = s;
    }
    ksum ks = {0, 0};
    for (long ij = 0; ij < nxy; ++ij) kadd(&ks, rows[ij]);
    q = ks.s + ks.c;
    free(rows);
  }
  free(rho2);
  return e + 0.5 * g * q * step_factor(nx, ny, nz, step);
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
