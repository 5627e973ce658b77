/*
 * Pass planner: maps a 3-D grid onto the pencil-pass kernels (grid_hip.h) and builds the
 * kinetic-phase tables.
 *
 * A grid is "squeezed" to its non-trivial axes, padded on the left with length-1 axes, so
 * the pass kernels always see (n0, n1, n2) with n2 contiguous.  The transform conventions
 * are the reference's (src/cgrid_fft.c:40-52 via FFTW): forward exp(-2 pi i ...), backward
 * exp(+2 pi i ...), both unnormalised, in place.
 *
 * Kinetic phase (reference src/grid_wf_fft.c:141-193): psi_k *= exp(tm * (kx^2+ky^2+kz^2)),
 * tm = -i t hbar / (2 m), k_a = 2 pi i / (n_a step) for i <= n_a/2 else 2 pi (i - n_a) /
 * (n_a step), minus k0_a.  The exponential of the sum is the product of one exponential per
 * axis, so the device applies three 1-D tables (computed here in extended precision and
 * rounded once) instead of evaluating a complex exponential per point.
 */
#include "internal.h"

void lg_squeeze(long nx, long ny, long nz, lg_shape *s) {
  long d[3] = {nx, ny, nz};
  int k = 2;
  s->n[0] = s->n[1] = s->n[2] = 1;
  s->phys[0] = s->phys[1] = s->phys[2] = -1;
  for (int a = 2; a >= 0; --a) {
    if (d[a] > 1) {
      s->n[k] = d[a];
      s->phys[k] = a;
      --k;
    }
  }
  s->m = 2 - k;
}

void lg_pass_init(hgrid_pass *p, const lg_shape *s, int prec, void *in, void *out, int axis) {
  memset(p, 0, sizeof(*p));
  p->in = in;
  p->out = out;
  p->prec = prec;
  p->n0 = s->n[0];
  p->n1 = s->n[1];
  p->n2 = s->n[2];
  p->axis = axis;
  p->nfft = 1;
  p->dir1 = -1;
  p->dir2 = +1;
  p->post = 1.0;
  p->mid = 1.0;
  /* FP64 twiddles for FP64 arithmetic (double grids and FP32 storage with FP64 arithmetic) */
  p->tw = lg_twiddles(s->n[axis], prec == HGRID_C64 ? HGRID_C64 : HGRID_C128);
}

/* ---- per-pass timing ------------------------------------------------------------------------
 * While enabled, every pass kernel is bracketed by two HIP events on the library stream and its
 * description kept, so a caller can attribute the time of a real step to its passes (the
 * counterpart of the reference's cuda_statistics, src/cuda.c). */
#define LG_PASS_TRACE_MAX 4096
typedef struct {
  void *ev[2];
  int axis, nfft, prec;
  unsigned flags;
  long len;
} pass_record;
static pass_record g_trace[LG_PASS_TRACE_MAX];
static int g_trace_on = 0, g_trace_n = 0, g_trace_events = 0;

void grid_hip_pass_timing(int enable) {
  if (enable && !g_trace_events) {
    lg_ensure_device();
    for (int i = 0; i < LG_PASS_TRACE_MAX; ++i)
      for (int k = 0; k < 2; ++k) lg_hip(hgrid_event_create(&g_trace[i].ev[k]), "hipEventCreate (pass timing)");
    g_trace_events = 1;
  }
  g_trace_on = enable ? 1 : 0;
  g_trace_n = 0;
}

INT grid_hip_pass_times(INT max, int *axis, int *nfft, unsigned *flags, long *len, int *prec, double *ms) {
  const INT n = g_trace_n < max ? g_trace_n : max;
  for (INT i = 0; i < n; ++i) {
    float t = 0.0f;
    lg_hip(hgrid_event_elapsed_ms(g_trace[i].ev[0], g_trace[i].ev[1], &t), "hipEventElapsedTime (pass timing)");
    axis[i] = g_trace[i].axis;
    nfft[i] = g_trace[i].nfft;
    flags[i] = g_trace[i].flags;
    len[i] = g_trace[i].len;
    prec[i] = g_trace[i].prec;
    ms[i] = t;
  }
  return n;
}

void lg_run_pass(hgrid_pass *p) {
  pass_record *r = (g_trace_on && g_trace_n < LG_PASS_TRACE_MAX) ? &g_trace[g_trace_n++] : NULL;
  if (r) lg_hip(hgrid_event_record(r->ev[0]), "hipEventRecord (pass timing)");
  int rc = hgrid_fft_pass(p);
  if (rc) lg_fatal("FFT pass (axis %d, shape %ld x %ld x %ld) failed: %s", p->axis, p->n0, p->n1, p->n2,
                   hgrid_last_error()[0] ? hgrid_last_error() : "unsupported configuration");
  if (r) {
    lg_hip(hgrid_event_record(r->ev[1]), "hipEventRecord (pass timing)");
    r->axis = p->axis;
    r->nfft = p->nfft;
    r->flags = p->flags;
    r->len = p->axis == 0 ? p->n0 : p->axis == 1 ? p->n1 : p->n2;
    r->prec = p->prec;
  }
}

static void copy_scaled(void *in, void *out, int prec, long n, double post, const double *scale) {
  size_t es = lg_prec_size(prec);
  if (in != out) lg_hip(hgrid_memcpy_d2d(out, in, (size_t)n * es), "hipMemcpy D2D");
  if (scale || post != 1.0) lg_hip(hgrid_cscale(out, prec, n, post, 0.0, scale), "cscale");
}

void lg_fft_forward(void *in, void *out, int prec, const lg_shape *s, const double *scale) {
  if (s->m == 0) {
    copy_scaled(in, out, prec, 1, 1.0, scale);
    return;
  }
  int first = 1;
  for (int axis = 2; axis >= 3 - s->m; --axis) {
    hgrid_pass p;
    lg_pass_init(&p, s, prec, first ? in : out, out, axis);
    p.dir1 = -1;
    if (first && scale) {
      p.flags |= HGRID_PF_PRESCALE;
      p.scale = scale;
    }
    lg_run_pass(&p);
    first = 0;
  }
}

void lg_fft_inverse(void *in, void *out, int prec, const lg_shape *s, double post, const double *scale) {
  if (s->m == 0) {
    copy_scaled(in, out, prec, 1, post, scale);
    return;
  }
  int first = 1;
  for (int axis = 3 - s->m; axis <= 2; ++axis) {
    hgrid_pass p;
    lg_pass_init(&p, s, prec, first ? in : out, out, axis);
    p.dir1 = +1;
    if (first && scale) {
      p.flags |= HGRID_PF_PRESCALE;
      p.scale = scale;
    }
    if (axis == 2) p.post = post;
    lg_run_pass(&p);
    first = 0;
  }
}

/* ---- kinetic tables ------------------------------------------------------------------------ */

static long double k_of(long i, long n, double step, double k0) {
  /* reference: kx = i * lx - kx0 (i <= n/2) else (i - n) * lx - kx0, lx = 2 pi / (step n) */
  const double l = 2.0 * M_PI / (step * (double)n);
  const double k = (i <= n / 2) ? ((double)i) * l - k0 : ((double)(i - n)) * l - k0;
  return (long double)k;
}

static void cexp_ld(long double ar, long double ai, long double *re, long double *im) {
  long double m = expl(ar);
  *re = m * cosl(ai);
  *im = m * sinl(ai);
}

/* Kinetic tables are FP64 (double2) for either storage precision: the kernels form the
 * factor in FP64 and round once onto FP32 data. */
static void upload_table(void **dptr, const long double *re, const long double *im, long n) {
  const size_t es = 16;
  double *h = (double *)malloc((size_t)n * es);
  if (!h) lg_fatal("out of host memory (tables)");
  for (long i = 0; i < n; ++i) {
    h[2 * i] = (double)re[i];
    h[2 * i + 1] = (double)im[i];
  }
  if (!*dptr) lg_hip(hgrid_malloc(dptr, (size_t)n * es), "hipMalloc (tables)");
  lg_hip(hgrid_memcpy_h2d(*dptr, h, (size_t)n * es), "hipMemcpy (tables)");
  free(h);
}

ktab_cache *lg_kinetic_tables(cgrid *g, double tm_re, double tm_im, double fold) {
  return lg_kinetic_tables_in(g, &g->dev->kin, tm_re, tm_im, fold);
}

/* the tables in a given cache slot of the grid (kin, or kin2 for a merged full step) */
ktab_cache *lg_kinetic_tables_in(cgrid *g, ktab_cache *k, double tm_re, double tm_im, double fold) {
  grid_dev *d = g->dev;
  if (k->valid && k->prec == d->prec && k->tm_re == tm_re && k->tm_im == tm_im && k->kx0 == g->kx0 && k->ky0 == g->ky0 &&
      k->kz0 == g->kz0 && k->step == g->step && k->fold == fold && k->nx == g->nx && k->ny == g->ny && k->nz == g->nz)
    return k;
  if (k->valid && (k->nx != g->nx || k->ny != g->ny || k->nz != g->nz || k->prec != d->prec)) {
    for (int a = 0; a < 3; ++a) {
      if (k->t[a]) {
        hgrid_free(k->t[a]);
        k->t[a] = NULL;
      }
      if (k->kc[a]) {
        hgrid_free(k->kc[a]);
        k->kc[a] = NULL;
      }
    }
  }
  long n[3] = {g->nx, g->ny, g->nz};
  double k0[3] = {g->kx0, g->ky0, g->kz0};
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  /* constant factor: fold times the factors of the trivial axes (k = -k0 there) */
  long double cr = (long double)fold, ci = 0.0L;
  for (int a = 0; a < 3; ++a) {
    if (n[a] != 1) continue;
    long double kk = k_of(0, 1, g->step, k0[a]);
    long double fr, fi;
    cexp_ld(tm_re * kk * kk, tm_im * kk * kk, &fr, &fi);
    long double nr = cr * fr - ci * fi, ni = cr * fi + ci * fr;
    cr = nr;
    ci = ni;
  }
  const int inner = s.m > 0 ? s.phys[2] : 2; /* the table that carries the constant factor */
  for (int a = 0; a < 3; ++a) {
    long len = n[a];
    long double *re = (long double *)malloc(sizeof(long double) * len), *im = (long double *)malloc(sizeof(long double) * len);
    if (!re || !im) lg_fatal("out of host memory (tables)");
    for (long i = 0; i < len; ++i) {
      long double kk = k_of(i, len, g->step, k0[a]);
      if (len == 1) {
        re[i] = 1.0L;
        im[i] = 0.0L;
      } else {
        cexp_ld(tm_re * kk * kk, tm_im * kk * kk, &re[i], &im[i]);
      }
      if (a == inner) {
        long double xr = re[i] * cr - im[i] * ci, xi = re[i] * ci + im[i] * cr;
        re[i] = xr;
        im[i] = xi;
      }
    }
    upload_table(&k->t[a], re, im, len);
    /* (k - k0)^2 as the reference's cutoff test forms it: k in FP64, squared in FP64
     * (src/grid_wf_fft.c:240-254); a trivial axis holds k0^2 */
    double *kc = (double *)malloc(sizeof(double) * len);
    if (!kc) lg_fatal("out of host memory (tables)");
    for (long i = 0; i < len; ++i) {
      const double kk = (double)k_of(i, len, g->step, k0[a]);
      kc[i] = kk * kk;
    }
    if (!k->kc[a]) lg_hip(hgrid_malloc((void **)&k->kc[a], sizeof(double) * len), "hipMalloc (cutoff tables)");
    lg_hip(hgrid_memcpy_h2d(k->kc[a], kc, sizeof(double) * len), "hipMemcpy (cutoff tables)");
    free(kc);
    free(re);
    free(im);
  }
  k->host_const_re = (double)cr;
  k->host_const_im = (double)ci;
  k->valid = 1;
  k->prec = d->prec;
  k->tm_re = tm_re;
  k->tm_im = tm_im;
  k->kx0 = g->kx0;
  k->ky0 = g->ky0;
  k->kz0 = g->kz0;
  k->step = g->step;
  k->fold = fold;
  k->nx = g->nx;
  k->ny = g->ny;
  k->nz = g->nz;
  return k;
}

ktab_cache *lg_k2_tables(cgrid *g) {
  grid_dev *d = g->dev;
  ktab_cache *k = &d->kin_energy;
  if (k->valid && k->step == g->step && k->nx == g->nx && k->ny == g->ny && k->nz == g->nz) return k;
  lg_ensure_device();
  long n[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; ++a) {
    double *h = (double *)malloc(sizeof(double) * n[a]);
    if (!h) lg_fatal("out of host memory (tables)");
    /* reference src/cgrid_diff.c:854-870: k = 2 pi i / (n step), i > n/2 -> i - n, no k0 */
    const double l = 2.0 * M_PI / (g->step * (double)n[a]);
    for (long i = 0; i < n[a]; ++i) {
      double kk = (i <= n[a] / 2) ? ((double)i) * l : ((double)(i - n[a])) * l;
      h[i] = kk * kk;
    }
    if (k->k2[a] && (k->nx != g->nx || k->ny != g->ny || k->nz != g->nz)) {
      hgrid_free(k->k2[a]);
      k->k2[a] = NULL;
    }
    if (!k->k2[a]) lg_hip(hgrid_malloc((void **)&k->k2[a], sizeof(double) * n[a]), "hipMalloc (k2 tables)");
    lg_hip(hgrid_memcpy_h2d(k->k2[a], h, sizeof(double) * n[a]), "hipMemcpy (k2 tables)");
    free(h);
  }
  k->valid = 1;
  k->step = g->step;
  k->nx = g->nx;
  k->ny = g->ny;
  k->nz = g->nz;
  return k;
}

int lg_sum_k2norm2(double *out, const void *psi, int prec, long nx, long ny, long nz, const double *k2x,
                   const double *k2y, const double *k2z, const double *scale) {
  long n = nx * ny * nz;
  double *scr = lg_partials(hgrid_reduce_scratch(n));
  return hgrid_sum_k2norm2(out, psi, prec, nx, ny, nz, k2x, k2y, k2z, scale, scr);
}
