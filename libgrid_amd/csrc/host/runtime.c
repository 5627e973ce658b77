/*
 * Library runtime: device bring-up, fatal errors, the shared reduction scratch and the
 * twiddle-table cache.  Plays the role of the reference's src/cuda.c (cuda_enable :1168,
 * cuda_status :1195, error checks :42-125) and src/threads.c / src/ctimer.c, for one GPU.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdarg.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "internal.h"

static int g_ready = 0;
static int g_device = 0;
static INT g_threads = 0;

void lg_fatal(const char *fmt, ...) {
  va_list ap;
  fprintf(stderr, "libgrid: ");
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  fflush(stderr);
  abort();
}

void lg_hip(int rc, const char *what) {
  if (rc) lg_fatal("%s failed: %s", what, hgrid_last_error());
}

void lg_ensure_device(void) {
  if (g_ready) return;
  int rc = hgrid_init(g_device);
  if (rc == -1) lg_fatal("no HIP device available (%s); this library has no CPU back end", hgrid_last_error());
  if (rc == -2) lg_fatal("%s", hgrid_last_error());
  if (rc) lg_fatal("device initialisation failed: %s", hgrid_last_error());
  g_ready = 1;
}

void grid_hip_enable(char value, int ngpus, int *gpus) {
  if (!value) lg_fatal("grid_hip_enable(0): the GPU cannot be disabled, this library has no CPU back end");
  if (ngpus > 1) lg_fatal("grid_hip_enable: %d GPUs requested; this build drives one GPU per process", ngpus);
  if (g_ready) {
    if (ngpus == 1 && gpus && gpus[0] != g_device) lg_fatal("grid_hip_enable: device already selected");
    return;
  }
  g_device = (ngpus == 1 && gpus) ? gpus[0] : 0;
  lg_ensure_device();
}

char grid_hip_status(void) { return (char)g_ready; }

void grid_hip_sync(void) {
  if (g_ready) lg_hip(hgrid_sync(), "hipStreamSynchronize");
}

size_t lg_prec_size(int prec) { return prec == HGRID_C128 ? 16 : 8; }

/* ---- reduction scratch ------------------------------------------------------------------ */

static double *g_partials = NULL;
static long g_partials_n = 0;
static unsigned long g_partials_gen = 0; /* counts reallocations: captured graphs hold the address */

double *lg_partials(long ndoubles) {
  lg_ensure_device();
  if (ndoubles > g_partials_n) {
    if (g_partials) lg_hip(hgrid_free(g_partials), "hipFree");
    long n = ndoubles < 65536 ? 65536 : ndoubles;
    lg_hip(hgrid_malloc((void **)&g_partials, (size_t)n * sizeof(double)), "hipMalloc (reduction scratch)");
    g_partials_n = n;
    ++g_partials_gen;
  }
  return g_partials;
}

unsigned long lg_partials_generation(void) { return g_partials_gen; }

/* ---- twiddle tables ----------------------------------------------------------------------- */

typedef struct {
  long n;
  int prec;
  void *d;
} tw_entry;
static tw_entry g_tw[64];
static int g_ntw = 0;

/* exp(-2 pi i m / n) evaluated in long double (x87 extended precision) and rounded once. */
static void twiddle_ld(long m, long n, double *re, double *im) {
  /* reduce to the first octant so the extended-precision sin/cos work on small arguments */
  long mm = m % n;
  long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)mm / (long double)n;
  *re = (double)cosl(a);
  *im = (double)(-sinl(a));
  /* exact values where they exist */
  if (mm == 0) { *re = 1.0; *im = 0.0; }
  else if (4 * mm == n) { *re = 0.0; *im = -1.0; }
  else if (2 * mm == n) { *re = -1.0; *im = 0.0; }
  else if (4 * mm == 3 * n) { *re = 0.0; *im = 1.0; }
}

const void *lg_twiddles(long n, int prec) {
  for (int i = 0; i < g_ntw; ++i)
    if (g_tw[i].n == n && g_tw[i].prec == prec) return g_tw[i].d;
  if (g_ntw == 64) lg_fatal("twiddle table cache full");
  lg_ensure_device();
  size_t es = lg_prec_size(prec);
  void *h = malloc((size_t)n * es);
  if (!h) lg_fatal("out of host memory (twiddles)");
  for (long m = 0; m < n; ++m) {
    double re, im;
    twiddle_ld(m, n, &re, &im);
    if (prec == HGRID_C128) {
      ((double *)h)[2 * m] = re;
      ((double *)h)[2 * m + 1] = im;
    } else {
      ((float *)h)[2 * m] = (float)re;
      ((float *)h)[2 * m + 1] = (float)im;
    }
  }
  void *d;
  lg_hip(hgrid_malloc(&d, (size_t)n * es), "hipMalloc (twiddles)");
  lg_hip(hgrid_memcpy_h2d(d, h, (size_t)n * es), "hipMemcpy (twiddles)");
  free(h);
  g_tw[g_ntw].n = n;
  g_tw[g_ntw].prec = prec;
  g_tw[g_ntw].d = d;
  ++g_ntw;
  return d;
}

/* ---- seeded random streams --------------------------------------------------------------- */

unsigned long long lg_splitmix64(unsigned long long *state) {
  unsigned long long z = (*state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

double lg_normal(unsigned long long *state) {
  /* two 53-bit uniforms in (0, 1] and (0, 1) */
  const double u1 = ((double)(lg_splitmix64(state) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(lg_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* ---- threads and timers (reference src/threads.c:89-157, src/ctimer.c:23-60) -------------- */

void grid_threads_init(INT threads) {
  if (threads < 0) lg_fatal("Error in grid_threads_init(). Number of threads <= 0.");
  if (threads == 0) threads = (INT)sysconf(_SC_NPROCESSORS_ONLN);
  if (g_threads && g_threads != threads)
    lg_fatal("Error in grid_threads_init(). This function was called twice with different number of threads.");
  g_threads = threads;
#ifdef _OPENMP
  omp_set_num_threads((int)threads);
#endif
  fprintf(stderr, "libgrid: Double precision floats, long integers, %ld host threads (gfx950 HIP back end).\n",
          (long)threads);
}

INT grid_threads(void) { return g_threads; }

void grid_timer_start(grid_timer *timer) {
  gettimeofday(&timer->zero_time, 0);
  timer->zero_clock = clock();
}

REAL grid_timer_wall_clock_time(grid_timer *timer) {
  struct timeval now;
  gettimeofday(&now, 0);
  return 1.0 * (REAL)(now.tv_sec - timer->zero_time.tv_sec) + 1e-6 * (REAL)(now.tv_usec - timer->zero_time.tv_usec);
}

REAL grid_timer_cpu_time(grid_timer *timer) {
  clock_t now = clock();
  return (REAL)(now - timer->zero_clock) / (REAL)CLOCKS_PER_SEC;
}
