/*
 * Batched 2-D imaginary-time ground states: many independent (1 x n x n) wavefunctions
 * relaxed in one persistent GPU launch (hip/batch2d.hpp), with complex64 storage and FP64
 * reductions, plus the autotuner that picks the kernel's tile shape.
 *
 * Per solve this is the reference composition (src/grid_wf.c:379-432 in imaginary time,
 * then grid_wf_normalize :694-700, with grid_wf_energy_fft src/grid_wf_fft.c:101-129 as
 * the convergence measure) repeated until the energy is converged to tol by the rule of the
 * host relaxation (relax_rule.h), with that relaxation's vector extrapolation.
 */
#include <sys/stat.h>
#include <unistd.h>

#include "internal.h"

typedef struct {
  int variant, nthreads, per_cu; /* tile variant, workgroup size, persistent workgroups per CU */
} b2d_tile;

static const b2d_tile kCandidates[] = {{0, 512, 1}, {0, 512, 2}, {0, 1024, 1}, {1, 512, 1}, {1, 512, 2}, {1, 1024, 1}};
#define NCAND ((int)(sizeof(kCandidates) / sizeof(kCandidates[0])))

static b2d_tile g_tile = {-1, 0, 0};
static INT g_cap = 0;               /* persistent-launch step cap (0: none, the default) */
static INT g_stride = 64;           /* vector-extrapolation stride inside the persistent launch (0: plain) */
static INT g_phase2_stride = 64;    /* vector-extrapolation stride of the straggler phase (0: plain) */
static INT g_last_extrap = 0;       /* extrapolations done inside the last launch, all solves */
static INT g_phase2_count = 0;      /* solves continued on the multi-kernel path by the last call */
static double g_phase2_seconds = 0.0;
static char g_report[2048];

static const char *cache_path(char *buf, size_t len) {
  const char *home = getenv("HOME");
  const char *xdg = getenv("XDG_CACHE_HOME");
  if (xdg && *xdg) snprintf(buf, len, "%s/libgrid_amd/batch2d_tile.txt", xdg);
  else snprintf(buf, len, "%s/.cache/libgrid_amd/batch2d_tile.txt", home ? home : "/tmp");
  return buf;
}

static int load_cached_tile(void) {
  char path[1024], dev[128];
  FILE *f = fopen(cache_path(path, sizeof(path)), "r");
  if (!f) return 0;
  b2d_tile t;
  int ok = fscanf(f, "%127s %d %d %d", dev, &t.variant, &t.nthreads, &t.per_cu) == 4;
  fclose(f);
  if (!ok || strcmp(dev, hgrid_device_name()) != 0) return 0;
  for (int i = 0; i < NCAND; ++i)
    if (kCandidates[i].variant == t.variant && kCandidates[i].nthreads == t.nthreads && kCandidates[i].per_cu == t.per_cu) {
      g_tile = t;
      return 1;
    }
  return 0;
}

static void save_cached_tile(void) {
  char path[1024];
  cache_path(path, sizeof(path));
  char dir[1024];
  snprintf(dir, sizeof(dir), "%s", path);
  char *slash = strrchr(dir, '/');
  if (slash) {
    *slash = 0;
    char *p = dir + 1;
    while ((p = strchr(p, '/'))) {
      *p = 0;
      mkdir(dir, 0755);
      *p++ = '/';
    }
    mkdir(dir, 0755);
  }
  FILE *f = fopen(path, "w");
  if (!f) return; /* a read-only home only costs a re-tune next time */
  fprintf(f, "%s %d %d %d\n", hgrid_device_name(), g_tile.variant, g_tile.nthreads, g_tile.per_cu);
  fclose(f);
}

/* ---- problem packing ---------------------------------------------------------------------- */

typedef struct {
  INT nsolves;
  hgrid_b2d_job *hjobs;
  hgrid_b2d_job *djobs;
  int *queue;
  double *out;
  float *psi_pool, *pv_pool;
  double *v_pool;
  float *snap_pool; /* 2 x float2 per point: extrapolation snapshot and difference (NULL: none) */
  void *tables; /* kh, kf, k2, tw per solve, packed */
  long *off;    /* element offset of each solve in the pools */
  float *host_psi0;
  long total;
  double *hist; /* per-step energies, hist_len per solve (input order) */
  long hist_len;
} b2d_batch;

/* energy history of the last relaxation (grid_batch2d_record_history) */
static long g_hist_len = 0;
static double *g_hist = NULL;
static INT g_hist_ns = 0;

static void b2d_free(b2d_batch *b) {
  hgrid_free(b->djobs);
  hgrid_free(b->queue);
  hgrid_free(b->out);
  hgrid_free(b->psi_pool);
  hgrid_free(b->pv_pool);
  hgrid_free(b->v_pool);
  if (b->snap_pool) hgrid_free(b->snap_pool);
  hgrid_free(b->tables);
  if (b->hist) hgrid_free(b->hist);
  free(b->hjobs);
  free(b->off);
  free(b->host_psi0);
}

static int cmp_desc(const void *a, const void *b) {
  const long *x = (const long *)a, *y = (const long *)b;
  return (x[0] < y[0]) - (x[0] > y[0]);
}

static void b2d_pack(b2d_batch *b, INT ns, wf **gwf, rgrid **pot, REAL dt, REAL tol, INT max_steps, INT stride) {
  memset(b, 0, sizeof(*b));
  lg_ensure_device();
  b->nsolves = ns;
  b->off = (long *)malloc(sizeof(long) * (ns + 1));
  long total = 0, tab_bytes = 0;
  for (INT i = 0; i < ns; ++i) {
    cgrid *g = gwf[i]->grid;
    if (g->nx != 1 || g->ny != g->nz || !hgrid_b2d_supported(g->ny))
      lg_fatal("batched 2-D solve %ld: grid must be 1 x n x n with n in {64, 96, 128, 192, 256}", (long)i);
    if (pot[i]->nx != 1 || pot[i]->ny != g->ny || pot[i]->nz != g->nz) lg_fatal("batched 2-D solve %ld: potential size", (long)i);
    if (gwf[i]->propagator != WF_2ND_ORDER_FFT) lg_fatal("batched 2-D solves need WF_2ND_ORDER_FFT");
    b->off[i] = total;
    total += g->ny * g->nz;
    tab_bytes += g->ny * (4 + 4 + 8 + 8);
  }
  b->off[ns] = total;
  b->total = total;
  float *hpsi = (float *)malloc(sizeof(float) * 2 * total), *hpv = (float *)malloc(sizeof(float) * total);
  double *hv = (double *)malloc(sizeof(double) * total);
  char *htab = (char *)malloc(tab_bytes);
  if (!hpsi || !hpv || !hv || !htab) lg_fatal("out of host memory (batched solves)");
  lg_hip(hgrid_malloc((void **)&b->psi_pool, sizeof(float) * 2 * total), "hipMalloc (batch psi)");
  lg_hip(hgrid_malloc((void **)&b->pv_pool, sizeof(float) * total), "hipMalloc (batch pv)");
  lg_hip(hgrid_malloc((void **)&b->v_pool, sizeof(double) * total), "hipMalloc (batch v)");
  lg_hip(hgrid_malloc(&b->tables, tab_bytes), "hipMalloc (batch tables)");
  if (stride > 0)
    lg_hip(hgrid_malloc((void **)&b->snap_pool, sizeof(float) * 4 * total), "hipMalloc (batch extrapolation snapshots)");
  lg_hip(hgrid_malloc((void **)&b->out, sizeof(double) * HGRID_B2D_NOUT * ns), "hipMalloc (batch out)");
  lg_hip(hgrid_malloc((void **)&b->queue, 256), "hipMalloc (batch queue)");
  b->hist_len = g_hist_len;
  if (b->hist_len > 0)
    lg_hip(hgrid_malloc((void **)&b->hist, sizeof(double) * (size_t)(b->hist_len * ns)), "hipMalloc (batch history)");
  b->hjobs = (hgrid_b2d_job *)calloc(ns, sizeof(hgrid_b2d_job));
  long toff = 0;
  /* queue order: largest grids first (longest solves start early) */
  long *order = (long *)malloc(sizeof(long) * 2 * ns);
  for (INT i = 0; i < ns; ++i) {
    order[2 * i] = gwf[i]->grid->ny * gwf[i]->grid->nz;
    order[2 * i + 1] = i;
  }
  qsort(order, ns, 2 * sizeof(long), cmp_desc);
  for (INT r = 0; r < ns; ++r) {
    INT i = order[2 * r + 1];
    cgrid *g = gwf[i]->grid;
    rgrid *v = pot[i];
    const long n = g->ny, nn = n * n, o = b->off[i];
    cgrid_host_sync(g);
    rgrid_host_sync(v);
    for (long k = 0; k < nn; ++k) {
      hpsi[2 * (o + k)] = (float)creal(g->value[k]);
      hpsi[2 * (o + k) + 1] = (float)cimag(g->value[k]);
      const long y = k / n, x = k % n;
      const double vv = v->value[y * v->nz2 + x];
      hv[o + k] = vv;
      hpv[o + k] = (float)exp(-dt * vv / HBAR);
    }
    /* tables: kh, kf (float), k2 (double), tw (float2); k_m with Nyquist at +n/2
     * (src/grid_wf_fft.c:173-187), factors and twiddles each rounded once */
    float *kh = (float *)(htab + toff);
    float *kf = (float *)(htab + toff + 4 * n);
    double *k2 = (double *)(htab + toff + 8 * n);
    float *tw = (float *)(htab + toff + 16 * n);
    const double l = 2.0 * M_PI / (g->step * (double)n);
    for (long m = 0; m < n; ++m) {
      const double k = (m <= n / 2) ? (double)m * l : (double)(m - n) * l;
      k2[m] = k * k;
      kh[m] = (float)exp(-0.5 * dt * HBAR * k * k / (2.0 * gwf[i]->mass));
      kf[m] = (float)exp(-dt * HBAR * k * k / (2.0 * gwf[i]->mass));
      const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
      tw[2 * m] = (float)cosl(a);
      tw[2 * m + 1] = (float)(-sinl(a));
    }
    hgrid_b2d_job *j = &b->hjobs[r];
    j->psi = b->psi_pool + 2 * o;
    j->pv = b->pv_pool + o;
    j->v = b->v_pool + o;
    j->kh = (const float *)((char *)b->tables + toff);
    j->kf = (const float *)((char *)b->tables + toff + 4 * n);
    j->k2 = (const double *)((char *)b->tables + toff + 8 * n);
    j->tw = (const void *)((char *)b->tables + toff + 16 * n);
    j->snap = b->snap_pool ? (void *)(b->snap_pool + 4 * o) : NULL;
    j->diff = b->snap_pool ? (void *)(b->snap_pool + 4 * o + 2 * nn) : NULL;
    j->stride = b->snap_pool ? (int)stride : 0;
    j->dA = g->step * g->step;
    /* E = (tfac sum k^2 |psi_k|^2 + sum V |psi|^2) / sum |psi|^2 for the kernel's unnormalised
     * transforms (the k-space sum of n^2 points carries n^2 |psi|^2 by Parseval, the real-space
     * sums n^4 |psi|^2 after the two round trips) */
    j->tfac = HBAR * HBAR / (2.0 * gwf[i]->mass) * (double)nn;
    j->tol = tol;
    j->n = (int)n;
    j->max_steps = (int)max_steps;
    j->out = b->out + HGRID_B2D_NOUT * i;
    j->hist_len = (int)b->hist_len;
    j->hist = b->hist ? b->hist + b->hist_len * i : NULL;
    toff += n * (4 + 4 + 8 + 8);
  }
  free(order);
  lg_hip(hgrid_memcpy_h2d(b->psi_pool, hpsi, sizeof(float) * 2 * total), "hipMemcpy (batch psi)");
  lg_hip(hgrid_memcpy_h2d(b->pv_pool, hpv, sizeof(float) * total), "hipMemcpy (batch pv)");
  lg_hip(hgrid_memcpy_h2d(b->v_pool, hv, sizeof(double) * total), "hipMemcpy (batch v)");
  lg_hip(hgrid_memcpy_h2d(b->tables, htab, tab_bytes), "hipMemcpy (batch tables)");
  lg_hip(hgrid_malloc((void **)&b->djobs, sizeof(hgrid_b2d_job) * ns), "hipMalloc (batch jobs)");
  lg_hip(hgrid_memcpy_h2d(b->djobs, b->hjobs, sizeof(hgrid_b2d_job) * ns), "hipMemcpy (batch jobs)");
  b->host_psi0 = hpsi;
  free(hpv);
  free(hv);
  free(htab);
}

static int b2d_blocks(const b2d_tile *t, INT ns) {
  int cus = hgrid_b2d_cus();
  if (cus <= 0) cus = 256;
  long nb = (long)t->per_cu * cus;
  if (nb > ns) nb = ns;
  return (int)nb;
}

static double b2d_run(b2d_batch *b, const b2d_tile *t) {
  void *e0, *e1;
  lg_hip(hgrid_event_create(&e0), "event");
  lg_hip(hgrid_event_create(&e1), "event");
  lg_hip(hgrid_event_record(e0), "event");
  int rc = hgrid_b2d_launch(b->djobs, (int)b->nsolves, b->queue, t->variant, t->nthreads, t->per_cu,
                            b2d_blocks(t, b->nsolves));
  if (rc) lg_fatal("batched 2-D launch failed: %s", hgrid_last_error());
  lg_hip(hgrid_event_record(e1), "event");
  float ms = 0.f;
  lg_hip(hgrid_event_elapsed_ms(e0, e1, &ms), "event");
  hgrid_event_destroy(e0);
  hgrid_event_destroy(e1);
  return ms;
}

/* Per-solve start/end of the last relaxation, in seconds from the earliest start (the device
 * clock read by the kernel runs at a constant 100 MHz). */
static double *g_t0 = NULL, *g_t1 = NULL;
static INT g_nt = 0;

static void solve_times(const double *out, INT ns) {
  free(g_t0);
  free(g_t1);
  g_t0 = (double *)malloc(sizeof(double) * ns);
  g_t1 = (double *)malloc(sizeof(double) * ns);
  if (!g_t0 || !g_t1) lg_fatal("out of host memory (batch times)");
  double first = 1e300;
  for (INT i = 0; i < ns; ++i)
    if (out[HGRID_B2D_NOUT * i + 3] < first) first = out[HGRID_B2D_NOUT * i + 3];
  for (INT i = 0; i < ns; ++i) {
    g_t0[i] = (out[HGRID_B2D_NOUT * i + 3] - first) * 1e-8;
    g_t1[i] = (out[HGRID_B2D_NOUT * i + 4] - first) * 1e-8;
  }
  g_nt = ns;
}

/* ---- public ------------------------------------------------------------------------------- */

void grid_batch2d_record_history(INT len) { g_hist_len = len > 0 ? len : 0; }

INT grid_batch2d_history(INT solve, REAL *energy, INT len) {
  if (!g_hist || solve < 0 || solve >= g_hist_ns) return 0;
  INT m = len < g_hist_len ? len : g_hist_len;
  for (INT k = 0; k < m; ++k) energy[k] = g_hist[g_hist_len * solve + k];
  return m;
}

INT grid_batch2d_solve_times(REAL *start, REAL *end) {
  for (INT i = 0; i < g_nt; ++i) {
    if (start) start[i] = g_t0[i];
    if (end) end[i] = g_t1[i];
  }
  return g_nt;
}

const char *grid_batch2d_autotune(INT nsolves, wf **gwf, rgrid **potential, REAL dt, INT probe_steps) {
  b2d_batch b;
  b2d_pack(&b, nsolves, gwf, potential, dt, 0.0 /* never converge */, probe_steps, g_stride);
  double best = 1e300;
  int bi = 0;
  size_t pos = (size_t)snprintf(g_report, sizeof(g_report), "batch2d autotune on %s (%ld solves, %ld steps):",
                                hgrid_device_name(), (long)nsolves, (long)probe_steps);
  for (int c = 0; c < NCAND; ++c) {
    lg_hip(hgrid_memcpy_h2d(b.psi_pool, b.host_psi0, sizeof(float) * 2 * b.total), "hipMemcpy (batch psi)");
    b2d_run(&b, &kCandidates[c]); /* warm-up (code load) */
    lg_hip(hgrid_memcpy_h2d(b.psi_pool, b.host_psi0, sizeof(float) * 2 * b.total), "hipMemcpy (batch psi)");
    double ms = b2d_run(&b, &kCandidates[c]);
    if (pos < sizeof(g_report))
      pos += (size_t)snprintf(g_report + pos, sizeof(g_report) - pos, " [v%d t%d x%d: %.3f ms]", kCandidates[c].variant,
                              kCandidates[c].nthreads, kCandidates[c].per_cu, ms);
    if (ms < best) {
      best = ms;
      bi = c;
    }
  }
  g_tile = kCandidates[bi];
  if (pos < sizeof(g_report))
    snprintf(g_report + pos, sizeof(g_report) - pos, " -> v%d t%d x%d", g_tile.variant, g_tile.nthreads, g_tile.per_cu);
  save_cached_tile();
  b2d_free(&b);
  return g_report;
}

void grid_batch2d_set_tile(int variant, int nthreads, int per_cu) {
  for (int i = 0; i < NCAND; ++i)
    if (kCandidates[i].variant == variant && kCandidates[i].nthreads == nthreads && kCandidates[i].per_cu == per_cu) {
      g_tile = kCandidates[i];
      return;
    }
  lg_fatal("grid_batch2d_set_tile: (%d, %d, %d) is not a tile configuration", variant, nthreads, per_cu);
}

void grid_batch2d_get_tile(int *variant, int *nthreads, int *per_cu) {
  *variant = g_tile.variant;
  *nthreads = g_tile.nthreads;
  *per_cu = g_tile.per_cu;
}

REAL grid_wf_batch2d_relax(INT nsolves, wf **gwf, rgrid **potential, REAL dt, REAL tol, INT max_steps, REAL *energy,
                           INT *steps) {
  if (nsolves <= 0) return 0.0;
  lg_ensure_device();
  if (g_tile.variant < 0 && !load_cached_tile()) {
    /* first use on this device: tune on (up to) the first 64 problems for 20 steps */
    grid_batch2d_autotune(nsolves < 64 ? nsolves : 64, gwf, potential, dt, 20);
  }
  /* The persistent launch relaxes every solve to convergence; nearly degenerate problems, whose
   * imaginary-time tails are long, are shortened by the launch's own vector extrapolation
   * (grid_batch2d_set_extrapolation).  With a step cap (grid_batch2d_set_straggler_cap, off by
   * default) solves unconverged at the cap continue one at a time on the multi-kernel FP32 path,
   * which puts the whole GPU on one grid. */
  const INT cap = (g_cap > 0 && g_cap < max_steps) ? g_cap : max_steps;
  b2d_batch b;
  b2d_pack(&b, nsolves, gwf, potential, dt, tol, cap, g_stride);
  const double ms = b2d_run(&b, &g_tile);
  double *out = (double *)malloc(sizeof(double) * HGRID_B2D_NOUT * nsolves);
  lg_hip(hgrid_memcpy_d2h(out, b.out, sizeof(double) * HGRID_B2D_NOUT * nsolves), "hipMemcpy (batch out)");
  solve_times(out, nsolves);
  g_last_extrap = 0;
  for (INT i = 0; i < nsolves; ++i) g_last_extrap += (INT)out[HGRID_B2D_NOUT * i + 5];
  if (b.hist) {
    free(g_hist);
    g_hist = (double *)malloc(sizeof(double) * (size_t)(b.hist_len * nsolves));
    if (!g_hist) lg_fatal("out of host memory (batch history)");
    lg_hip(hgrid_memcpy_d2h(g_hist, b.hist, sizeof(double) * (size_t)(b.hist_len * nsolves)), "hipMemcpy (history)");
    g_hist_ns = nsolves;
  }
  float *hpsi = (float *)malloc(sizeof(float) * 2 * b.total);
  lg_hip(hgrid_memcpy_d2h(hpsi, b.psi_pool, sizeof(float) * 2 * b.total), "hipMemcpy (batch psi)");
  for (INT i = 0; i < nsolves; ++i) {
    cgrid *g = gwf[i]->grid;
    const long nn = g->ny * g->nz, o = b.off[i];
    for (long k = 0; k < nn; ++k) g->value[k] = (double)hpsi[2 * (o + k)] + I * (double)hpsi[2 * (o + k) + 1];
    cgrid_host_modified(g);
    if (out[HGRID_B2D_NOUT * i + 2] == -2.0 || !isfinite(out[HGRID_B2D_NOUT * i]))
      lg_fatal("batched relaxation: solve %ld (%ld x %ld) diverged: non-finite energy", (long)i, (long)g->ny, (long)g->nz);
    if (energy) energy[i] = out[HGRID_B2D_NOUT * i];
    if (steps) steps[i] = (INT)out[HGRID_B2D_NOUT * i + 2];
  }
  b2d_free(&b);
  free(hpsi);
  /* phase 2 */
  grid_timer tm;
  grid_timer_start(&tm);
  g_phase2_count = 0;
  for (INT i = 0; i < nsolves && cap < max_steps; ++i) {
    if ((INT)out[HGRID_B2D_NOUT * i + 2] < cap) continue;
    cgrid *g = gwf[i]->grid;
    const char prec0 = cgrid_get_precision(g);
    cgrid_set_precision(g, GRID_PREC_SINGLE);
    REAL e2 = 0.0;
    const INT s2 = g_phase2_stride
                       ? grid_wf_gpe_relax_extrapolated(gwf[i], potential[i], 0.0, dt, tol, max_steps - cap, &e2,
                                                        g_phase2_stride, NULL)
                       : grid_wf_gpe_relax(gwf[i], potential[i], 0.0, dt, tol, max_steps - cap, &e2);
    cgrid_set_precision(g, prec0);
    if (energy) energy[i] = e2;
    if (steps) steps[i] = cap + s2;
    ++g_phase2_count;
  }
  grid_hip_sync();
  g_phase2_seconds = g_phase2_count ? grid_timer_wall_clock_time(&tm) : 0.0;
  free(out);
  return ms * 1e-3 + g_phase2_seconds;
}

void grid_batch2d_set_straggler_cap(INT steps) { g_cap = steps; }

void grid_batch2d_set_extrapolation(INT stride) {
  if (stride < 0) lg_fatal("grid_batch2d_set_extrapolation: stride %ld is negative", (long)stride);
  g_stride = stride;
}

INT grid_batch2d_last_extrapolations(void) { return g_last_extrap; }

void grid_batch2d_set_straggler_extrapolation(INT stride) {
  if (stride < 0 || stride % 16) lg_fatal("grid_batch2d_set_straggler_extrapolation: stride %ld is not a multiple of 16", (long)stride);
  g_phase2_stride = stride;
}

void grid_batch2d_last_phase2(INT *count, REAL *seconds) {
  if (count) *count = g_phase2_count;
  if (seconds) *seconds = g_phase2_seconds;
}
