/*
 * Thin C-ABI between libgrid's C host code and its hand-written HIP kernels (gfx950).
 *
 * Everything here takes plain device pointers, sizes and scalars; no C99 complex, no HIP
 * types, no torch types.  The C host library (grid.h API) is the only intended caller; the
 * functions are exported from libgrid.so so that tests and the benchmark can time them.
 *
 * All launches go to one HIP stream per process (created by hgrid_init at default priority);
 * every call below is asynchronous unless it says otherwise.  A non-zero return value means
 * the HIP runtime reported an error; hgrid_last_error() describes it.
 *
 * It replaces the reference's "char hook -> extern C ...W launcher" seam
 * (src/grid_wf-cuda.c:18-49 -> src/grid_wf-cuda2.cu:72-104, src/cgrid_fft-cuda.c:41-174,
 * src/cgrid-cuda2.cu:1187-1220), but with device-canonical grids and fused FFT passes.
 */

#ifndef LIBGRID_AMD_GRID_HIP_H
#define LIBGRID_AMD_GRID_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---------------------------------------------------------------------------- */

/* Select the device and create the library stream.  Returns 0 on success, -1 if no HIP device
 * is present, -2 if the device is not gfx950 (the code objects are built for gfx950 only). */
int hgrid_init(int device);
int hgrid_initialized(void);
const char *hgrid_last_error(void);
const char *hgrid_device_name(void);          /* gcnArchName of the selected device */
int hgrid_malloc(void **ptr, size_t bytes);
int hgrid_free(void *ptr);
int hgrid_memcpy_h2d(void *dst, const void *src, size_t bytes);   /* synchronous */
int hgrid_memcpy_d2h(void *dst, const void *src, size_t bytes);   /* synchronous */
int hgrid_memcpy_d2d(void *dst, const void *src, size_t bytes);   /* async, library stream */
int hgrid_memset(void *ptr, int value, size_t bytes);             /* async fill kernel */
int hgrid_copy_f64(double *dst, const double *src, long n);        /* async copy kernel (graph-safe) */
int hgrid_sync(void);                                              /* wait for the library stream */
/* Elapsed time of the library stream between two marks (HIP events recorded on that stream). */
int hgrid_event_create(void **ev);
int hgrid_event_destroy(void *ev);
int hgrid_event_record(void *ev);
int hgrid_event_elapsed_ms(void *start, void *stop, float *ms);   /* synchronizes on stop */
int hgrid_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Streaming copy of `bytes` (a multiple of 16) with 16-byte loads/stores: the measured HBM
 * ceiling the FFT passes are compared with, and the calibration kernel for PMC byte counts. */
int hgrid_stream_copy(void *dst, const void *src, size_t bytes);

/* ---- storage conversion ----------------------------------------------------------------- */

/* FP64 complex <-> FP32 complex (device arrays of n elements).  Reduced-precision storage
 * keeps a host complex128 mirror; these convert on upload / download. */
int hgrid_c128_to_c64(void *dst64, const void *src128, size_t n);
int hgrid_c64_to_c128(void *dst128, const void *src64, size_t n);
/* Pack an nz2-padded real rgrid row layout into a dense [nx*ny][nz] array (and back). */
int hgrid_rpack(double *dst, const double *src, long rows, long nz, long nz2);
int hgrid_runpack(double *dst, const double *src, long rows, long nz, long nz2);

/* ---- FFT passes ------------------------------------------------------------------------- */

/* Precision of a device complex array. */
#define HGRID_C128 0   /* double2 */
#define HGRID_C64  1   /* float2 (twiddles and phase tables are computed in FP64, stored FP32) */
#define HGRID_C64X 2   /* float2 storage; the pass kernels widen to double2 and do all arithmetic in FP64
                          (FP64 twiddles); other kernels treat it as float2 */

/* Pass flags (bit set).  Stage order inside one pass:
 *   load [* *scale] -> PRE ops -> FFT #1 -> MID ops -> [FFT #2] -> * post -> POST ops -> store */
#define HGRID_PF_PRESCALE   0x0001u  /* multiply loads by the complex device scalar *scale */
#define HGRID_PF_PRE_RHO    0x0002u  /* pre: rho_out = |psi|^2 (GPE density of the input state) */
#define HGRID_PF_PRE_RED    0x0004u  /* pre: partial sums of |psi|^2, vext|psi|^2, |psi|^4 */
#define HGRID_PF_MID_KIN    0x0008u  /* mid: psi *= ktab_a[n] * ktab_1[c1] * ktab_2[c2] (double2 tables) */
#define HGRID_PF_MID_KRED   0x0010u  /* mid: partial sum of k^2 |psi_k|^2 (before MID_KIN) */
#define HGRID_PF_MID_POT    0x0020u  /* mid: psi *= mid * exp(c * (cons + pot)), complex pot */
#define HGRID_PF_MID_GPE    0x0040u  /* mid: psi *= mid * exp(c * (cons + vext + g * rho)),
                                        * exp(c * (cons + rho)) when vext is NULL (rho holds V_eff),
                                        * or exp(c * (cons + vext)) when rho is NULL */
#define HGRID_PF_POST_RED   0x0080u  /* post: partial sum of |psi|^2 of the stored values */
#define HGRID_PF_NOSTORE    0x0100u  /* reduce only; do not write the array */
#define HGRID_PF_PRE_VEFF   0x0200u  /* pre: rho_out = vext + g |psi|^2 (GPE potential) instead of |psi|^2 */
#define HGRID_PF_PRE_AFTER  0x0800u  /* the PRE_RHO / PRE_VEFF / PRE_RED operations act on the transformed
                                        data (a real-space state stored transformed along this axis) */
#define HGRID_PF_KIN_REAL   0x0400u  /* with MID_KIN: the tables are real (imaginary time); the factor is
                                        applied as an FP64 real scale */
#define HGRID_PF_KCUT       0x1000u  /* with MID_KIN: zero psi where (kc_a[n] + kc_1[c1]) + kc_2[c2] >= kcut2
                                        (the k-space cutoff of the CFFT propagator) */
#define HGRID_PF_POT_NOFFT1 0x2000u  /* with MID_POT: no transform before the potential factor
                                        (factor, then transform dir2); nfft = 1 */
#define HGRID_PF_POT_NOFFT2 0x4000u  /* with MID_POT: no transform after it (transform dir1, then
                                        factor); nfft = 1.  MID_POT passes also take PRESCALE */

/* Reduction channels (partials[chan * nblocks + block]). */
#define HGRID_RED_NORM   0   /* sum |psi|^2 of the pass input   (PRE_RED) */
#define HGRID_RED_VEXT   1   /* sum vext |psi|^2                 (PRE_RED, needs vext) */
#define HGRID_RED_QUART  2   /* sum |psi|^4                      (PRE_RED) */
#define HGRID_RED_KIN    3   /* sum k^2 |psi_k|^2                (MID_KRED) */
#define HGRID_RED_POST   4   /* sum |psi|^2 of the pass output   (POST_RED) */
#define HGRID_RED_NCHAN  5

typedef struct hgrid_pass {
  void *in;              /* device array (complex, precision prec) */
  void *out;             /* may equal in (in-place) */
  int prec;              /* HGRID_C128 / HGRID_C64 / HGRID_C64X */
  long n0, n1, n2;       /* squeezed shape; n2 is contiguous */
  int axis;              /* 0, 1 or 2 */
  int nfft;              /* 1 or 2 transforms along the axis in this pass */
  int dir1, dir2;        /* -1 forward (exp(-i..)), +1 inverse (exp(+i..)), unnormalised */
  unsigned flags;
  const void *tw;        /* W_L^m = exp(-2 pi i m / L), m < L, precision prec, L = length of axis */
  const double *scale;   /* PRESCALE: device complex scalar {re, im} */
  double post;           /* factor applied to the stored values (1.0 = none) */
  double mid;            /* factor applied at MID_POT / MID_GPE */
  /* kinetic tables (precision prec): factor for the element index along axis, and for the two
   * other squeezed coordinates c1, c2 of the pencil (see hgrid_pass_coords) */
  const void *ktab_a, *ktab_1, *ktab_2;
  const double *k2_a, *k2_1, *k2_2;   /* |k|^2 per axis for MID_KRED (no k0 shift); the kernels
                                       * form the pass axis's own entries from k2_dk */
  const double *kc_a, *kc_1, *kc_2;   /* (k - k0)^2 per axis for KCUT */
  double kcut2;                       /* KCUT threshold kmax^2 */
  /* potential */
  const void *pot;       /* complex potential, same layout and precision as in */
  const double *vext;    /* dense real external potential (FP64) */
  const double *rho;     /* dense real density (FP64) for MID_GPE */
  double g;              /* nonlinear coupling */
  double c_re, c_im;     /* c = -i * tstep / hbar */
  double cons;           /* constant added to the potential */
  double *rho_out;       /* PRE_RHO destination (FP64, dense) */
  double *partials;      /* HGRID_RED_NCHAN * nblocks doubles (see hgrid_pass_blocks) */
  double k2_dk;          /* MID_KRED: k spacing 2 pi / (step L) of the pass axis; element i has
                          * k = i dk (i <= L/2) or (i - L) dk and |k|^2 = k * k, as in k2_a */
} hgrid_pass;

/* Number of workgroups a pass launches (size of each partials channel). */
long hgrid_pass_blocks(int prec, long n0, long n1, long n2, int axis);
/* Launch one pass. Returns 0, or -1 for an unsupported axis length / precision. */
int hgrid_fft_pass(const hgrid_pass *p);
/* Axis lengths the pass kernels support (power of two, 1..1024). */
int hgrid_fft_supported(long n);

/* ---- reductions ------------------------------------------------------------------------- */

/* out[c] = sum_{b < nblocks} partials[c * nblocks + b] for c < nchan, deterministic order,
 * FP64 Neumaier-compensated.  out is a device array. */
int hgrid_reduce_partials(double *out, const double *partials, long nblocks, int nchan);

/* Up to HGRID_MAX_JOBS reductions of pass partials in one launch (the end of a fused step):
 * job j sums channels [chan_first, chan_first + nchan) of its partials into out[0..nchan).
 * Then, if scale != NULL: with target > 0, *scale = {sqrt(target / (*norm_sum * dv)), 0}
 * (renormalisation); with target <= 0, *scale = {1, 0} (a consumed pending scale). */
#define HGRID_MAX_JOBS 4
typedef struct hgrid_reduce_job {
  const double *partials;
  long nblocks;
  int chan_first, nchan;
  double *out;
} hgrid_reduce_job;
typedef struct hgrid_finalize {
  hgrid_reduce_job job[HGRID_MAX_JOBS];
  int njobs;
  double *scale;
  const double *norm_sum;
  double target, dv;
} hgrid_finalize;
int hgrid_step_finalize(const hgrid_finalize *f);

/* Elementwise reductions over device arrays (n elements), written to the device scalar *out
 * (deterministic; scratch must hold hgrid_reduce_scratch(n) doubles). */
long hgrid_reduce_scratch(long n);
int hgrid_sum_norm2(double *out, const void *a, int prec, long n, const double *scale, double *scratch);
/* hgrid_sum_norm2 without scratch: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).  It allocates
 * nothing: the partials go to a fixed 128 KiB array per device (2 * 8192 + 2 doubles, enough for
 * every n, since a reduction uses at most 8192 workgroups), looked up once at hgrid_init.  n must
 * be in 1 .. 2^34, else it returns -1 with an error and launches nothing.  Every call on a device
 * shares that array, so calls that overlap on one device (other streams, other threads) race;
 * calls on the library stream are ordered. */
int hgrid_sum_norm2_noalloc(double *out, const void *a, int prec, long n, const double *scale);
int hgrid_sum_vnorm2(double *out, const void *a, int prec, const double *v, long n, const double *scale, double *scratch);
int hgrid_sum_conj_product(double *out2, const void *a, const void *b, int prec, long n, double *scratch);
/* sum (k2x[i] + k2y[j] + k2z[k]) |a[i,j,k]|^2 over an nx x ny x nz array */
int hgrid_sum_k2norm2(double *out, const void *a, int prec, long nx, long ny, long nz, const double *k2x,
                      const double *k2y, const double *k2z, const double *scale, double *scratch);
/* sum a[i] (squared = 0) or sum a[i]^2 (squared = 1) of a dense real array */
int hgrid_sum_real(double *out, const double *a, long n, int squared, double *scratch);

/* ---- pointwise kernels ------------------------------------------------------------------ */

/* psi *= exp(c * (cons + pot)) with c = (c_re, c_im); pot complex (same precision). */
int hgrid_potential_phase(void *psi, const void *pot, int prec, long n, double c_re, double c_im, double cons,
                          const double *scale);
/* psi[i,j,k] *= ta[i] * tb[j] * tc[k]  (3 tables of double2 factors for either precision;
 * the factor is formed in FP64 and rounded once onto psi).  With kcx != NULL, psi[i,j,k] = 0
 * where (kcx[i] + kcy[j]) + kcz[k] >= kcut2 (the CFFT cutoff; kc* hold (k - k0)^2). */
int hgrid_kinetic_phase(void *psi, int prec, long nx, long ny, long nz, const void *ta, const void *tb,
                        const void *tc, const double *scale, const double *kcx, const double *kcy,
                        const double *kcz, double kcut2);
/* dst = (1 - alpha) wr + alpha wi with the absorbing-boundary amplitude alpha(i, j, k) of the
 * box bounds = {lx, hx, ly, hy, lz, hz} (host array); dst may be wr */
int hgrid_wf_merge(void *dst, const void *wr, const void *wi, int prec, long nx, long ny, long nz, const long *bounds);
/* c += (dr + i di) a, n elements of precision prec */
int hgrid_caxpy(void *c, const void *a, int prec, long n, double dr, double di);
/* psi *= (c_re + i c_im) * (*scale if scale) */
int hgrid_cscale(void *psi, int prec, long n, double c_re, double c_im, const double *scale);
/* dst = |src|^2 (dense FP64 output) */
int hgrid_density(double *dst, const void *src, int prec, long n, const double *scale);
/* dst = (complex) src (real part, imaginary zero) */
int hgrid_real_to_complex(void *dst, int prec, const double *src, long n);
/* dst = c (complex constant) */
int hgrid_cfill(void *dst, int prec, long n, double c_re, double c_im);
/* device complex scalar ops on a 2-double slot: s = {re, im} */
int hgrid_scalar_set(double *s, double re, double im);
int hgrid_scalar_mul(double *s, double re, double im);
/* s = sqrt(target / (sum[0] * dv)) (renormalisation factor from a device sum) */
int hgrid_scalar_renorm(double *s, const double *sum, double target, double dv);
/* FD kinetic energy integrand: out = sum Re(conj(psi) * laplace7(psi)) / h^2 (periodic,
 * Neumann or Dirichlet boundary via bc = 2/1/0; a Dirichlet neighbour outside the grid has the
 * value dir_re + i dir_im, which the pending scale does not multiply). */
int hgrid_fd_laplace_expect(double *out, const void *psi, int prec, long nx, long ny, long nz, int bc,
                            double dir_re, double dir_im, const double *scale, double *scratch);

/* HIP graphs of the library stream (no memset nodes are ever recorded: the library clears
 * device memory with kernels) */
int hgrid_graph_begin(void);
int hgrid_graph_end(void **exec);
int hgrid_graph_launch(void *exec);
int hgrid_graph_free(void *exec);

/* ---- batched 2-D imaginary-time solves ---------------------------------------------------- */

/* One n x n solve of the persistent batched kernel (complex64 storage, FP64 reductions). */
typedef struct hgrid_b2d_job {
  void *psi;           /* n*n float2, row-major [y][x]; in: start state, out: normalised ground state */
  const float *pv;     /* n*n exp(-dt V) */
  const double *v;     /* n*n V (potential energy) */
  const float *kh;     /* n: exp(-(dt/2) hbar k_m^2 / 2m), the same table for both axes */
  const float *kf;     /* n: exp(-dt hbar k_m^2 / 2m) */
  const double *k2;    /* n: k_m^2 */
  const void *tw;      /* n: float2 exp(-2 pi i m / n), rounded once */
  void *snap, *diff;   /* n*n float2 each: vector-extrapolation snapshot and difference (stride > 0) */
  double dA;           /* step^2 */
  double tfac;         /* hbar^2/(2m) * n^2 (energy = (tfac sum k^2|psi_k|^2 + sum V|psi|^2) / sum |psi|^2
                          with the unnormalised transforms of the kernel) */
  double tol;          /* energy tolerance of the stopping rule (see grid_wf_gpe_relax) */
  int n, max_steps;
  int stride;          /* vector-extrapolation stride in steps (0: none) */
  int hist_len;        /* energies entering the first hist_len steps go to hist (NaN between the
                          stopping rule's checkpoints, where no energy is formed; 0: none) */
  double *hist;
  double *out;         /* device, HGRID_B2D_NOUT doubles: energy, steps taken, step at which tol
                          was met (max_steps if never, -2 if the energy became non-finite), start
                          and end of the solve on the device's constant 100 MHz clock,
                          extrapolations done */
} hgrid_b2d_job;
#define HGRID_B2D_NOUT 6

int hgrid_b2d_supported(long n);   /* n in {64, 96, 128, 192, 256} */
int hgrid_b2d_cus(void);            /* compute units of the device */
/* jobs: device array; queue: device int (reset by the call).  Tile choices: variant 0/1
 * (threads per pencil for n = 64 and 256), nthreads 512 or 1024 per workgroup, per_cu
 * workgroups per CU (2 only with 512 threads; sets the register budget), nblocks workgroups
 * (persistent: each takes solves from the queue until it is empty). */
int hgrid_b2d_launch(const hgrid_b2d_job *jobs, int njobs, int *queue, int variant, int nthreads, int per_cu,
                     int nblocks);

/* ---- warm-start network kernels (BF16 activations, FP32 accumulation) ---------------------
 * No counterpart in the reference.  Activations are channels-last [voxel][channel] BF16 with a
 * channel stride (s) and offset (o) per tensor, so concatenations are slices of one buffer.
 * Voxels m = ((b S + z) S + y) S + x over cubes of side S = 2^lgS; 3x3x3 convolutions wrap
 * periodically within each sample.  Channel counts are multiples of 8 (16 for outputs). */

/* y[m][yo + co] = act(sum_t sum_ci x[m + d_t][xo + ci] w[co][t][ci] + bias[co]) (bias may be
 * NULL; act = ReLU if relu) and then zero where mask[m][mko + co] <= 0 (mask may be NULL).
 * With wt from hgrid_nn_wprep in place of w it is the data gradient of the convolution. */
int hgrid_nn_conv3(const void *x, int xs, int xo, int cin, const void *w, const float *bias, void *y, int ys, int yo,
                   int cout, const void *mask, int mks, int mko, int lgS, long nvox, int relu);
/* partial[p][co][t][ci] of dW = sum_m dy[m][dyo + co] x[m + d_t][xo + ci]; sum the
 * hgrid_nn_wgrad_partials(nvox) partials with hgrid_nn_reduce.  bpartial (may be NULL)
 * receives [chunk][co] column sums of dy (the bias gradient), hgrid_nn_wgrad_chunks(nvox)
 * rows. */
int hgrid_nn_conv3_wgrad(const void *dy, int dys, int dyo, int cout, const void *x, int xs, int xo, int cin, int lgS,
                         long nvox, float *partial, float *bpartial);
long hgrid_nn_wgrad_chunks(long nvox);
long hgrid_nn_wgrad_partials(long nvox);
long hgrid_nn_wgrad_chunk(long nvox);
int hgrid_nn_reduce(const float *partial, long np, long n, float *out, int accumulate);
int hgrid_nn_reduce_strided(const float *partial, long np, int stride, int ncols, float *out);
/* partial[chunk][c] = column sums of d[.][dof + c]; hgrid_nn_colsum_chunks(nvox) chunks */
int hgrid_nn_colsum(const void *d, int ds, int dof, int c, long nvox, float *partial);
long hgrid_nn_colsum_chunks(long nvox);
/* nearest upsampling to the fine level (side 2^lgF) and its gradient times (act > 0) */
int hgrid_nn_up2(const void *src, int ss, int so, int c, void *dst, int ds, int dof, int lgF, long nfine);
int hgrid_nn_up2_bwd(const void *dfine, int fs, int fo, int c, const void *act, int as, int ao, void *dst, int lgC,
                     long ncoarse);
/* 2x2x2 average pooling to a dense coarse tensor, and the gradient of a pooled activation plus
 * its skip gradient (dskip may be NULL) times (act > 0), written dense */
int hgrid_nn_pool2(const void *src, int ss, int so, int c, void *dst, int lgC, long ncoarse);
int hgrid_nn_pool2_bwd(const void *dcoarse, const void *dskip, int ks, int ko, int c, const void *act, int as, int ao,
                       void *dst, int lgF, long nfine);
/* 1x1 head over 16 channels (wh: 16 weights then the bias) and its backward for the loss
 * sum (out - target)^2 * scale / 2: writes the masked input gradient and per-block partials
 * [block][18] (16 weight gradients, bias gradient, squared error) */
int hgrid_nn_head_fwd(const void *d, const float *wh, float *out, long nvox);
int hgrid_nn_head_bwd(const void *d, const float *wh, const float *out, const float *target, float scale, void *dpre,
                      long nvox, float *partial);
long hgrid_nn_head_blocks(long nvox);
/* BF16 weights wb[co][t][ci] and reflected-transposed wt[ci][26 - t][co] from FP32 w */
int hgrid_nn_wprep(const float *w, int cout, int cin, void *wb, void *wt);
/* Adam step; the gradient is multiplied by gscale[0] (may be NULL), and 0 there skips the step */
int hgrid_nn_adam(float *w, const float *g, float *m1, float *m2, long n, float lr, float b1, float b2, float eps, int t,
                  const float *gscale);
/* factor[0] = min(1, clip / |g|) (0 if |g| is not finite), factor[1] = |g|; scratch: 512 doubles */
int hgrid_nn_clip_factor(const float *g, long n, double clip, double *scratch, float *factor);
/* input tensor [m][8]: vscale * v[m], then zeros */
int hgrid_nn_input(const float *v, float vscale, void *x0, long nvox);
/* one S^3 sample under a symmetry of the periodic cube (axis permutation perm < 6, reflections
 * in the bits of flips, then a periodic shift) */
int hgrid_nn_symgather(const float *src, float *dst, int lgS, int perm, int flips, int s0, int s1, int s2);
/* periodic trilinear interpolation of an S^3 field onto n^3 (n = fct S) as a real wavefunction */
int hgrid_nn_to_grid(const float *a, int lgS, int fct, long n, void *psi, int prec);

#ifdef __cplusplus
}
#endif

#endif
