/*
 * libgrid (MI355X / gfx950 build): public C API for the FFT split-operator hot path.
 *
 * Drop-in for the reference library's header for the functions below: same names, argument
 * order and meaning, same in-place, unnormalised FFT conventions and the same struct field
 * names that callers read (value, nx/ny/nz, step, x0..z0, kx0..kz0, fft_norm, fft_norm2 /
 * grid, mass, norm, propagator, boundary).  Compile-time configuration matches the
 * reference's default build: REAL = double, INT = long, REAL complex = double _Complex
 * (reference make.conf:13-14, src/structs.h:24-103).
 *
 * Execution model: every computational call runs on the GPU.  A grid's device copy is
 * canonical while it exists; host accessors (cgrid_value_at_index, cgrid_write, ...) copy it
 * back first.  Code that reads or writes grid->value directly must call cgrid_host_sync()
 * (read) or cgrid_host_modified() (after writing).  There is no CPU back end: without a
 * gfx950 device every computational call prints "libgrid: ..." and aborts.
 */

#ifndef LIBGRID_AMD_GRID_H
#define LIBGRID_AMD_GRID_H

#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>

#ifdef __cplusplus
#error "grid.h is a C99 header (it uses C99 complex types)"
#endif

/* macros, not typedefs, so that `REAL complex` is a valid type (reference make.conf:62-87) */
#define REAL double
#define INT long

#ifndef HBAR
#define HBAR 1.0 /* reference make.conf:43-44 */
#endif

#define FMT_R "%le"
#define FMT_I "%ld"

#define CEXP cexp
#define CREAL creal
#define CIMAG cimag
#define CONJ conj
#define SQRT sqrt
#define EXP exp
#define COS cos
#define SIN sin
#define POW pow

#define GRID_EPS 1.0E-16

/* Wavefunction boundary conditions (reference src/defs.h:21-23) */
#define WF_DIRICHLET_BOUNDARY 1
#define WF_NEUMANN_BOUNDARY 2
#define WF_PERIODIC_BOUNDARY 3

/* Wavefunction propagators (reference src/defs.h:26-32).  This build implements the FFT ones
 * (1-4): WF_2ND_ORDER_FFT, WF_4TH_ORDER_FFT, and their CFFT forms, which remove the components
 * |k - k0| >= kmax in the kinetic phase.  grid_wf_alloc accepts the others and
 * grid_wf_propagate aborts for them. */
#define WF_2ND_ORDER_FFT 1
#define WF_4TH_ORDER_FFT 2
#define WF_2ND_ORDER_CFFT 3
#define WF_4TH_ORDER_CFFT 4
#define WF_2ND_ORDER_CN 5
#define WF_4TH_ORDER_CN 6
#define WF_1ST_ORDER_EULER 7

/* Device storage precision of a complex grid (this build's addition) */
#define GRID_PREC_DOUBLE 0 /* complex128 on the device (default) */
#define GRID_PREC_SINGLE 1 /* complex64 on the device; FP64 twiddles, phase arguments, reductions */
#define GRID_PREC_SINGLE_STORAGE 2 /* complex64 on the device; the FFT/propagation passes do all their
                                      arithmetic in FP64 and round only on store */

struct grid_dev;

typedef struct grid_timer_struct {
  struct timeval zero_time;
  clock_t zero_clock;
} grid_timer;

/* Complex grid (reference src/structs.h:24-50).  The FFTW/cuFFT plan fields are absent; the
 * device state is kept behind `dev`. */
typedef struct cgrid_struct {
  REAL complex *value;       /* host copy of the values, [nx][ny][nz], z fastest */
  char id[32];
  size_t grid_len;           /* bytes of value */
  INT nx, ny, nz;
  REAL step;
  REAL x0, y0, z0;           /* real space origin */
  REAL kx0, ky0, kz0;        /* reciprocal space origin (moving frame) */
  REAL omega;                /* rotation frequency; must be 0 for the FFT propagator */
  REAL complex (*value_outside)(struct cgrid_struct *grid, INT i, INT j, INT k);
  void *outside_params_ptr;
  REAL complex default_outside_params;
  char host_lock;            /* kept for source compatibility; must stay 0 (no CPU back end) */
  REAL fft_norm;             /* 1 / (nx ny nz) */
  REAL fft_norm2;            /* fft_norm * step^d, d = number of dimensions > 1 */
  char flag;                 /* claim / release */
  struct grid_dev *dev;      /* device residency (owned by the library) */
} cgrid;

/* Real grid (reference src/structs.h:57-83); host rows are padded to nz2 = 2 (nz/2 + 1). */
typedef struct rgrid_struct {
  REAL *value;
  char id[32];
  size_t grid_len;
  INT nx, ny, nz, nz2;
  REAL step;
  REAL x0, y0, z0;
  REAL kx0, ky0, kz0;
  REAL (*value_outside)(struct rgrid_struct *grid, INT i, INT j, INT k);
  void *outside_params_ptr;
  REAL default_outside_params;
  char host_lock;
  REAL fft_norm;
  REAL fft_norm2;
  char flag;
  struct grid_dev *dev;
} rgrid;

/* Wavefunction (reference src/structs.h:90-103) */
typedef struct wf_struct {
  cgrid *grid;
  REAL mass;
  REAL norm;                 /* target norm for grid_wf_normalize (default 1) */
  REAL cfft_width;
  char boundary;
  char propagator;
  cgrid *cworkspace, *cworkspace2, *cworkspace3;
  REAL (*ts_func)(INT, INT, INT, INT, INT, INT, INT, INT, INT);
  INT lx, hx, ly, hy, lz, hz;
  REAL kmax;
} wf;

/* ---- boundary functions (reference src/cgrid_bc.c:21-86, src/rgrid_bc.c) ---------------- */
REAL complex cgrid_value_outside_dirichlet(cgrid *grid, INT i, INT j, INT k);
REAL complex cgrid_value_outside_neumann(cgrid *grid, INT i, INT j, INT k);
REAL complex cgrid_value_outside_periodic(cgrid *grid, INT i, INT j, INT k);
REAL rgrid_value_outside_dirichlet(rgrid *grid, INT i, INT j, INT k);
REAL rgrid_value_outside_neumann(rgrid *grid, INT i, INT j, INT k);
REAL rgrid_value_outside_periodic(rgrid *grid, INT i, INT j, INT k);
#define CGRID_DIRICHLET_BOUNDARY cgrid_value_outside_dirichlet
#define CGRID_NEUMANN_BOUNDARY cgrid_value_outside_neumann
#define CGRID_PERIODIC_BOUNDARY cgrid_value_outside_periodic
#define RGRID_DIRICHLET_BOUNDARY rgrid_value_outside_dirichlet
#define RGRID_NEUMANN_BOUNDARY rgrid_value_outside_neumann
#define RGRID_PERIODIC_BOUNDARY rgrid_value_outside_periodic

/* ---- runtime ------------------------------------------------------------------------------ */
void grid_threads_init(INT threads);                      /* reference src/threads.c:89 */
INT grid_threads(void);                                   /* reference src/threads.c:150 */
void grid_timer_start(grid_timer *timer);                 /* reference src/ctimer.c:23 */
REAL grid_timer_wall_clock_time(grid_timer *timer);       /* reference src/ctimer.c:40 */
REAL grid_timer_cpu_time(grid_timer *timer);              /* reference src/ctimer.c:55 */
/* GPU on/off; replaces cuda_enable(char, int, int*) (reference src/cuda.c:1168).  One GPU:
 * ngpus must be 1 (or 0 = default device).  Disabling is refused: there is no CPU back end. */
void grid_hip_enable(char value, int ngpus, int *gpus);
char grid_hip_status(void);                               /* replaces cuda_status (src/cuda.c:1195) */
void grid_hip_sync(void);                                 /* wait for all queued GPU work */

/* ---- complex grids (reference src/cgrid.c) --------------------------------------------- */
cgrid *cgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL complex (*value_outside)(cgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id);                         /* :45 */
cgrid *cgrid_clone(cgrid *grid, char *id);                                      /* :169 */
void cgrid_free(cgrid *grid);                                                   /* :332 */
void cgrid_claim(cgrid *grid);                                                  /* :220 */
void cgrid_release(cgrid *grid);                                                /* :238 */
void cgrid_set_origin(cgrid *grid, REAL x0, REAL y0, REAL z0);                 /* :262 */
void cgrid_shift_origin(cgrid *grid, REAL x0, REAL y0, REAL z0);               /* :280 */
void cgrid_set_momentum(cgrid *grid, REAL kx0, REAL ky0, REAL kz0);            /* :302 */
void cgrid_set_rotation(cgrid *grid, REAL omega);                              /* :318 */
/* binary grid file [INT nx, ny, nz][REAL step][nx*ny*nz complex values]; a file of another size
 * or step is interpolated onto the grid (cgrid_extrapolate) */
void cgrid_write(cgrid *grid, FILE *out);                                       /* :360 */
cgrid *cgrid_read(cgrid *grid, FILE *in);                                       /* :382 */
/* base.grd (binary, as cgrid_write) and the ASCII cuts base.x/.y/.z through the middle point */
void cgrid_write_grid(char *base, cgrid *grid);                                 /* :432 */
/* value(i, j, k) = func(farg, (i - nx/2) step - x0, (j - ny/2) step - y0, (k - nz/2) step - z0) */
void cgrid_map(cgrid *grid, REAL complex (*func)(void *arg, REAL x, REAL y, REAL z), void *farg); /* :793 */
/* the same, each value a hat-weighted average of ns^3 samples over the neighbouring cells */
void cgrid_smooth_map(cgrid *grid, REAL complex (*func)(void *arg, REAL x, REAL y, REAL z), void *farg,
                      INT ns);                                                  /* :832 */
REAL complex cgrid_value(cgrid *grid, REAL x, REAL y, REAL z);                  /* :2060, trilinear */
void cgrid_extrapolate(cgrid *dest, cgrid *src);                                /* :2126 */
void cgrid_zero(cgrid *grid);
void cgrid_constant(cgrid *grid, REAL complex c);
void cgrid_copy(cgrid *dst, cgrid *src);
void cgrid_multiply(cgrid *grid, REAL complex c);                              /* :1223 */
void cgrid_add_scaled(cgrid *gridc, REAL complex d, cgrid *grida);              /* :1303 */
REAL cgrid_integral_of_square(cgrid *grid);                                     /* :1610 */
REAL complex cgrid_integral_of_conjugate_product(cgrid *grida, cgrid *gridb);   /* :1640 */
void cgrid_fft(cgrid *grid);                                                    /* :1822 */
void cgrid_inverse_fft(cgrid *grid);                                            /* :1839 */
void cgrid_scaled_inverse_fft(cgrid *grid, REAL complex c);                     /* :1857 */
void cgrid_inverse_fft_norm(cgrid *grid);                                       /* :1873 */
void cgrid_inverse_fft_norm2(cgrid *grid);                                      /* :1887 */
REAL complex cgrid_value_at_index(cgrid *grid, INT i, INT j, INT k);            /* :1960 */
void cgrid_value_to_index(cgrid *grid, INT i, INT j, INT k, REAL complex value);
REAL cgrid_fft_laplace_expectation_value(cgrid *grid);                          /* src/cgrid_diff.c:817 */
/* device storage precision (GRID_PREC_DOUBLE / GRID_PREC_SINGLE / GRID_PREC_SINGLE_STORAGE); this
 * build's addition */
void cgrid_set_precision(cgrid *grid, char prec);
char cgrid_get_precision(cgrid *grid);
/* host/device coherence for code that touches grid->value directly (this build's addition) */
void cgrid_host_sync(cgrid *grid);
void cgrid_host_modified(cgrid *grid);

/* ---- real grids (reference src/rgrid.c) ------------------------------------------------ */
rgrid *rgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL (*value_outside)(rgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id);                         /* :38 */
void rgrid_free(rgrid *grid);
void rgrid_map(rgrid *grid, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg); /* :806 */
void rgrid_smooth_map(rgrid *grid, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg, INT ns); /* :846 */
/* binary grid file [INT nx, ny, nz][REAL step][nx*ny*nz2 values, the padded rows] */
void rgrid_write(rgrid *grid, FILE *out);                                       /* :339 */
rgrid *rgrid_read(rgrid *grid, FILE *in);                                       /* :360 */
void rgrid_write_grid(char *base, rgrid *grid);                                 /* :505 */
REAL rgrid_value(rgrid *grid, REAL x, REAL y, REAL z);                          /* :2208, trilinear */
void rgrid_extrapolate(rgrid *dest, rgrid *src);                                /* :2267 */
void rgrid_zero(rgrid *grid);
void rgrid_constant(rgrid *grid, REAL c);
REAL rgrid_value_at_index(rgrid *grid, INT i, INT j, INT k);
void rgrid_value_to_index(rgrid *grid, INT i, INT j, INT k, REAL value);
REAL rgrid_integral(rgrid *grid);
REAL rgrid_integral_of_square(rgrid *grid);
void rgrid_host_sync(rgrid *grid);
void rgrid_host_modified(rgrid *grid);

/* ---- mixed (reference src/ngrid.c) ----------------------------------------------------- */
REAL grid_grid_expectation_value(cgrid *dgrid, rgrid *opgrid);                  /* :355 */
void grid_real_to_complex_re(cgrid *dest, rgrid *source);                       /* :53 */

/* ---- wavefunctions (reference src/grid_wf.c, src/grid_wf_fft.c, src/grid_wf_cn.c) ------- */
wf *grid_wf_alloc(INT nx, INT ny, INT nz, REAL step, REAL mass, char boundary, char propagator, char *id); /* :26 */
wf *grid_wf_clone(wf *gwf, char *id);                                           /* :131 */
void grid_wf_free(wf *gwf);                                                     /* :157 */
void grid_wf_propagate(wf *gwf, cgrid *potential, REAL complex time);          /* :379 */
void grid_wf_propagate_potential(wf *gwf, REAL complex tstep, cgrid *potential, REAL cons); /* :569 */
void grid_wf_propagate_kinetic_fft(wf *gwf, REAL complex time);                /* grid_wf_fft.c:141 */
void grid_wf_propagate_kinetic_cfft(wf *gwf, REAL complex time);               /* grid_wf_fft.c:206 */
/* Absorbing boundary for the 2nd-order FFT propagators: outside [lx,hx] x [ly,hy] x [lz,hz] a
 * real-time step is blended with an imaginary-time one by grid_wf_absorb's amplitude.  All
 * bounds 0 clears it. */
void grid_wf_boundary(wf *gwf, wf *gwfp, INT lx, INT hx, INT ly, INT hy, INT lz, INT hz); /* :99 */
REAL grid_wf_absorb(INT i, INT j, INT k, INT lx, INT hx, INT ly, INT hy, INT lz, INT hz); /* :185 */
void grid_wf_merge(wf *dst, cgrid *wfr, cgrid *wfi);                           /* :728 */
void grid_wf_density(wf *gwf, rgrid *density);                                  /* :611 */
void grid_wf_zero(wf *gwf);
void grid_wf_constant(wf *gwf, REAL complex c);
void grid_wf_map(wf *gwf, REAL complex (*func)(void *arg, REAL x, REAL y, REAL z), void *farg);
REAL grid_wf_norm(wf *gwf);                                                     /* :681 */
REAL grid_wf_normalize(wf *gwf);                                                /* :694 */
REAL complex grid_wf_overlap(wf *gwfa, wf *gwfb);
REAL grid_wf_energy(wf *gwf, rgrid *potential);                                 /* :215 */
REAL grid_wf_kinetic_energy(wf *gwf);                                           /* :230 */
REAL grid_wf_potential_energy(wf *gwf, rgrid *potential);                       /* :246 */
REAL grid_wf_energy_fft(wf *gwf, rgrid *potential);                             /* grid_wf_fft.c:101 */
REAL grid_wf_kinetic_energy_fft(wf *gwf);                                       /* grid_wf_fft.c:119 */
REAL grid_wf_energy_cn(wf *gwf, rgrid *potential);                              /* grid_wf_cn.c:19 */
REAL grid_wf_kinetic_energy_cn(wf *gwf);                                        /* grid_wf_cn.c:37 */
void grid_wf_analyze_method(char method);                                       /* grid_wf_analyze.c:21 */

/* ---- Gross-Pitaevskii driver (this build's addition; composition of reference calls) ----
 * One 2nd-order FFT step of i dpsi/dt = (T + vext + g|psi|^2) psi with the nonlinear
 * potential built from the state at the start of the step, exactly as
 *   grid_wf_density -> (g * rho + vext) -> grid_real_to_complex_re -> grid_wf_propagate
 * would, followed by grid_wf_normalize when normalize != 0 — fused into 8 or 9 FFT passes.  It
 * also records E_GP = T + int vext|psi|^2 + (g/2) int |psi|^4 of the start-of-step state. */
void grid_wf_gpe_propagate(wf *gwf, rgrid *vext, REAL g, REAL complex time, char normalize);
/* E_GP of the state that started the most recent grid_wf_gpe_propagate call. */
REAL grid_wf_gpe_last_energy(wf *gwf);
/* E_GP of the current state (FFT kinetic energy, no k0 shift, as grid_wf_kinetic_energy_fft). */
REAL grid_wf_gpe_energy(wf *gwf, rgrid *vext, REAL g);
/* Imaginary-time relaxation with renormalisation every step until the energy is converged to
 * tol, or max_steps.  Converged: with energies sampled every m steps (m = 16, doubling up to
 * 4096 while the decay is slow) the last decrease D2 is below tol and so is the geometric
 * estimate D2 q / (1 - q), q = D2 / D1, of the energy still to come, with two successive
 * ratios q within a factor of two; a rise smaller than tol (rounding) also ends it.  Steps run
 * in batches of 16 and only the first step of a batch, where every checkpoint falls, forms the
 * energy.  Returns the number of steps applied to the state (a multiple of 16 unless max_steps
 * ends it) and stores the energy of the returned state in *energy (may be NULL).  A non-finite
 * energy aborts with a libgrid: message (the relaxation diverged). */
INT grid_wf_gpe_relax(wf *gwf, rgrid *vext, REAL g, REAL dt, REAL tol, INT max_steps, REAL *energy);
/* grid_wf_gpe_relax with vector extrapolation (this build's addition; no reference
 * counterpart): every `stride` steps (a multiple of 16) a snapshot is taken, and when the
 * differences of three successive snapshots are parallel, d2 = r d1 with 0.05 < r < 0.995, the
 * geometric tail d2 r / (1 - r) of that slowest mode is added and the state renormalised.  The
 * stopping rule restarts after each extrapolation, so convergence is still decided by plain
 * steps.  Uses three grids of workspace.  *n_extrap (may be NULL) receives the number of
 * extrapolations. */
INT grid_wf_gpe_relax_extrapolated(wf *gwf, rgrid *vext, REAL g, REAL dt, REAL tol, INT max_steps, REAL *energy,
                                   INT stride, INT *n_extrap);
/* Relaxation replays batches of 16 steps as a HIP graph after the first batch (default on);
 * grid_hip_graphs(0) launches every step eagerly instead. */
void grid_hip_graphs(int enable);
/* Linear relaxations (g = 0) of 3-D grids with WF_2ND_ORDER_FFT merge the last kinetic half-step
 * of each step with the first of the next inside a batch (default on): four passes per step
 * instead of seven, the same states at the batch boundaries.  grid_hip_merged_relaxation(0)
 * runs whole steps. */
void grid_hip_merged_relaxation(int enable);
/* Per-pass timing (this build's addition; plays the role of the reference's cuda_statistics,
 * src/cuda.c): grid_hip_pass_timing(1) starts recording every FFT/propagation pass kernel
 * between two HIP events on the library stream (up to 4096 passes; enabling again resets the
 * record), 0 stops it.  grid_hip_pass_times copies up to max records: the pass's axis (0 outer,
 * 2 contiguous), number of transforms, flags (HGRID_PF_*), length, precision and its time in
 * milliseconds; it waits for those kernels.  Not for use under graph capture. */
void grid_hip_pass_timing(int enable);
INT grid_hip_pass_times(INT max, int *axis, int *nfft, unsigned *flags, long *len, int *prec, double *ms);

/* ---- batched 2-D ground states (this build's addition) ----------------------------------
 * Relax nsolves independent 1 x n x n wavefunctions (n in {64, 96, 128, 192, 256}) in
 * imaginary time with their potentials (rgrid, same shape) in one persistent GPU launch:
 * complex64 storage, FP64 twiddle arguments / tables / reductions.  Each solve runs 2nd-order
 * FFT steps of length dt, renormalised every step, until its FFT energy is converged to tol by
 * the rule of grid_wf_gpe_relax (or max_steps), with the vector extrapolation of
 * grid_wf_gpe_relax_extrapolated inside the launch.  The normalised states are written back
 * to gwf[i]->grid; energy[i] and steps[i] (may be NULL) receive the energy of that state and
 * the step at which tol was met (max_steps if never).  Returns the GPU time of the launch
 * (plus the wall time of the optional straggler phase) in seconds.  The first call on a
 * device runs grid_batch2d_autotune unless a tuned tile is cached. */
REAL grid_wf_batch2d_relax(INT nsolves, wf **gwf, rgrid **potential, REAL dt, REAL tol, INT max_steps, REAL *energy,
                           INT *steps);
/* Vector-extrapolation stride of the persistent launch in steps (default 64; 0 relaxes plainly). */
void grid_batch2d_set_extrapolation(INT stride);
/* Extrapolations the last grid_wf_batch2d_relax launch made, summed over its solves. */
INT grid_batch2d_last_extrapolations(void);
/* Optional straggler phase (off by default, 0): solves still unconverged after this many steps
 * of the persistent launch are continued one at a time on the multi-kernel FP32 path, which
 * spreads one grid over the whole GPU; grid_batch2d_last_phase2 reports how many and their
 * wall time. */
void grid_batch2d_set_straggler_cap(INT steps);
/* The straggler phase relaxes with vector extrapolation every stride steps
 * (grid_wf_gpe_relax_extrapolated; default 64); 0 relaxes the stragglers plainly. */
void grid_batch2d_set_straggler_extrapolation(INT stride);
void grid_batch2d_last_phase2(INT *count, REAL *seconds);
/* Time every tile configuration of the batched kernel on these problems for probe_steps steps,
 * keep the fastest (also cached in $XDG_CACHE_HOME or ~/.cache/libgrid_amd/batch2d_tile.txt)
 * and return a one-line report. */
const char *grid_batch2d_autotune(INT nsolves, wf **gwf, rgrid **potential, REAL dt, INT probe_steps);
/* Start and end (seconds from the earliest start, device clock) of every solve of the last
 * grid_wf_batch2d_relax call, in its input order; returns the count.  Arrays may be NULL. */
INT grid_batch2d_solve_times(REAL *start, REAL *end);
/* Record the energy entering each of the first len steps of every solve in later
 * grid_wf_batch2d_relax calls (0 turns it off; NaN on steps between the stopping rule's
 * checkpoints, where the launch forms no energy); grid_batch2d_history copies solve's record
 * of the last call (at most len values, entries past its last step undefined) and returns the
 * number copied. */
void grid_batch2d_record_history(INT len);
INT grid_batch2d_history(INT solve, REAL *energy, INT len);
void grid_batch2d_set_tile(int variant, int nthreads, int per_cu);
void grid_batch2d_get_tile(int *variant, int *nthreads, int *per_cu);

/* ---- learned warm start (this build's addition; no counterpart in the reference) --------
 * A 3-D convolutional U-Net (about 1.45 M parameters, BF16 weights and activations, FP32
 * accumulation on the matrix cores) maps a potential sampled on an S^3 grid to a ground-state
 * amplitude, which grid_wf_warm_start interpolates onto the wavefunction's n^3 grid (n a
 * multiple of S) as the start of imaginary-time relaxation.  The relaxation alone decides the
 * converged energy; with no network the plain Gaussian start is used. */
typedef struct grid_surrogate grid_surrogate;
/* side S: power of two >= 16; batch: samples per forward/backward; vscale: input scale of V */
grid_surrogate *grid_surrogate_alloc(INT side, INT batch, REAL vscale, INT seed);
void grid_surrogate_free(grid_surrogate *s);
INT grid_surrogate_nparams(grid_surrogate *s);
INT grid_surrogate_side(grid_surrogate *s);
/* parameters in layer order e1 e2a e2b e3a e3b ba bb d3a d3b d2a d2b d1 (W [Cout][27][Cin]
 * then b [Cout] each), then the head (16 weights, 1 bias); setting them resets Adam */
void grid_surrogate_set_params(grid_surrogate *s, const REAL *p);
void grid_surrogate_get_params(grid_surrogate *s, REAL *p);
/* loss (mean squared error) and its gradient (may be NULL) on nb <= batch samples [nb][S^3] */
REAL grid_surrogate_gradient(grid_surrogate *s, INT nb, const REAL *pot, const REAL *target, REAL *grad);
REAL grid_surrogate_loss(grid_surrogate *s, INT n, const REAL *pot, const REAL *target);
void grid_surrogate_predict(grid_surrogate *s, INT n, const REAL *pot, REAL *amp);
/* copy n samples to the device, then train for epochs (Adam, cosine-decayed lr) */
void grid_surrogate_set_training_data(grid_surrogate *s, INT n, const REAL *pot, const REAL *target);
REAL grid_surrogate_train(grid_surrogate *s, INT epochs, REAL lr, INT seed, REAL *loss_history);
/* as grid_surrogate_train, with flags: GRID_SURROGATE_AUGMENT draws every sample under a
 * random symmetry of the periodic cube (48 point operations and the translations);
 * GRID_SURROGATE_PROGRESS prints one line per epoch to stderr */
#define GRID_SURROGATE_AUGMENT 1
#define GRID_SURROGATE_PROGRESS 2
REAL grid_surrogate_train_ex(grid_surrogate *s, INT epochs, REAL lr, INT seed, int flags, REAL *loss_history);
int grid_surrogate_save(grid_surrogate *s, const char *path);
grid_surrogate *grid_surrogate_load(const char *path, INT batch);
/* starts: the network's (s may be NULL: Gaussian of fallback_width) and the plain Gaussian
 * exp(-r^2 / (2 width^2)) at the grid centre; both normalised */
void grid_wf_warm_start(wf *gwf, rgrid *potential, grid_surrogate *s, REAL fallback_width);
void grid_wf_gaussian_start(wf *gwf, REAL width);

#endif
