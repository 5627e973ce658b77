/*
 * Internal state of the C host library: device residency of grids, library scratch, the
 * FFT/phase table caches and the pass planner.  Not part of the public API.
 */
#ifndef LIBGRID_AMD_INTERNAL_H
#define LIBGRID_AMD_INTERNAL_H

#include "grid.h"
#include "grid_hip.h"

/* validity bits of a grid's two copies */
#define GV_HOST 1
#define GV_DEV 2

/* device scalar slots of one grid (doubles) */
#define GS_SCALE 0   /* pending complex scale {re, im} */
#define GS_RED 2     /* HGRID_RED_NCHAN reduction results of the last fused step */
#define GS_TMP 8     /* results of standalone reductions */
#define GS_N 16

typedef struct ktab_cache {
  int valid;
  int prec;
  double tm_re, tm_im;     /* time_mass = -i t hbar / (2m) */
  double kx0, ky0, kz0, step, fold;
  long nx, ny, nz;
  void *t[3];              /* device complex factor tables per physical axis */
  double *k2[3];           /* device |k|^2 tables per physical axis (no k0 shift) */
  double *kc[3];           /* device (k - k0)^2 tables per physical axis (CFFT cutoff) */
  double host_const_re, host_const_im;  /* product of trivial-axis factors (already folded) */
} ktab_cache;

typedef struct grid_dev {
  int kind;                /* 0 complex, 1 real */
  int prec;                /* HGRID_C128 / HGRID_C64 (complex grids) */
  size_t n;                /* number of elements (dense) */
  void *d;                 /* device array, or NULL */
  int valid;               /* GV_HOST | GV_DEV */
  int pending;             /* device scale slot differs from 1 */
  int zpend;               /* the device array holds the unnormalised forward transform of the
                              state along the squeezed contiguous axis (fused steps leave it so
                              between steps); accessors transform back before other use */
  double *dscal;           /* GS_N device doubles */
  /* wavefunction helpers (complex grids used as wf->grid) */
  void *work;              /* device-only scratch of the same size (kinetic energy etc.) */
  double *rho;             /* device density scratch for the GPE step */
  ktab_cache kin;          /* kinetic tables for the last (time, mass, frame) */
  ktab_cache kin2;         /* the full-step tables of a merged relaxation step (twice that time) */
  int kpend;               /* inside a merged relaxation batch: the array holds the state before
                              the last half kinetic factor of the step (never seen by callers) */
  ktab_cache kin_energy;   /* |k|^2 tables for energy reductions */
  unsigned long stamp;     /* bumped whenever the device array may have been written */
  void *alt;               /* the data converted to another storage precision (lg_c_in_as) */
  int alt_prec;
  unsigned long alt_stamp; /* stamp of the data alt was converted from */
  double last_energy_parts[4];
  int have_gpe_energy;
} grid_dev;

/* runtime.c */
void lg_fatal(const char *fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));
void lg_ensure_device(void);
void lg_hip(int rc, const char *what);          /* abort with the HIP error text if rc != 0 */
double *lg_partials(long ndoubles);             /* library reduction scratch (device) */
unsigned long lg_partials_generation(void);     /* changes whenever lg_partials reallocates */
const void *lg_twiddles(long n, int prec);      /* device table exp(-2 pi i m / n), m < n */
size_t lg_prec_size(int prec);

/* residency (cgrid.c / rgrid.c) */
grid_dev *lg_cdev(cgrid *g);                    /* allocate device state if needed */
void *lg_c_in(cgrid *g, const double **scale);  /* device data for reading; scale = pending or NULL */
void *lg_c_in_flushed(cgrid *g);                /* device data with any pending scale applied */
/* device data with any pending scale applied, in storage precision prec: the grid's own array,
 * or a converted copy kept beside it (the grid keeps its own precision) */
const void *lg_c_in_as(cgrid *g, int prec);
void *lg_c_out(cgrid *g);                       /* device data to be overwritten (no upload) */
void *lg_c_inout(cgrid *g, const double **scale);
void lg_c_written(cgrid *g, int scale_consumed);
void lg_c_set_pending(cgrid *g, double re, double im); /* multiply the pending scale */
void lg_c_flush(cgrid *g);
void *lg_c_raw(cgrid *g, const double **scale);  /* device data as stored (zpend kept), for the fused step */
void lg_c_materialize(cgrid *g);                 /* undo zpend: inverse transform along the contiguous axis */
void lg_c_host_sync(cgrid *g);
grid_dev *lg_rdev(rgrid *g);
const double *lg_r_in(rgrid *g);
double *lg_r_out(rgrid *g);
void lg_r_written(rgrid *g);
double lg_step_factor(INT nx, INT ny, INT nz, REAL step); /* step^d for dims != 1 */
unsigned long long lg_splitmix64(unsigned long long *state); /* seeded stream (SplitMix64) */
double lg_normal(unsigned long long *state);                 /* N(0, 1) by Box-Muller */

/* planner (fftplan.c) */
typedef struct lg_shape {
  long n[3];      /* squeezed shape n0, n1, n2 (n2 contiguous) */
  int phys[3];    /* physical axis (0 x, 1 y, 2 z) of each squeezed axis, -1 = padding */
  int m;          /* number of non-trivial axes */
} lg_shape;
void lg_squeeze(long nx, long ny, long nz, lg_shape *s);
void lg_pass_init(hgrid_pass *p, const lg_shape *s, int prec, void *in, void *out, int axis);
void lg_run_pass(hgrid_pass *p);
/* forward / inverse 3-D transform of one device array through the pass kernels */
void lg_fft_forward(void *in, void *out, int prec, const lg_shape *s, const double *scale);
void lg_fft_inverse(void *in, void *out, int prec, const lg_shape *s, double post, const double *scale);
/* kinetic tables: exp(tm (k - k0)^2) per physical axis; `fold` multiplies the table of the
 * innermost non-trivial axis (and absorbs trivial axes' constant factors). */
ktab_cache *lg_kinetic_tables(cgrid *g, double tm_re, double tm_im, double fold);
ktab_cache *lg_kinetic_tables_in(cgrid *g, ktab_cache *k, double tm_re, double tm_im, double fold);
ktab_cache *lg_k2_tables(cgrid *g);
/* sum over an array of (k2x[i] + k2y[j] + k2z[k]) |psi|^2 -> device out */
int lg_sum_k2norm2(double *out, const void *psi, int prec, long nx, long ny, long nz, const double *k2x,
                   const double *k2y, const double *k2z, const double *scale);

#endif
