/*
 * Complex grids: allocation, host/device residency, host-side mapping and element access,
 * FFT wrappers and reductions.  Follows the reference's src/cgrid.c (alloc :45-158, free
 * :332-349, write/read :360-420, map :793-815, multiply :1223-1238, integral_of_square
 * :1610-1629, FFT wrappers :1822-1890, value_at_index :1960-2000) and src/cgrid_bc.c.
 *
 * The device copy is canonical once created; the host array is refreshed on demand.  A
 * grid can carry a pending complex scale (a device scalar): cgrid_multiply and
 * grid_wf_normalize only update that scalar and the next kernel that reads the grid folds
 * it into its loads.
 */
#include "internal.h"

static void cdev_init(cgrid *g) {
  grid_dev *d = (grid_dev *)calloc(1, sizeof(grid_dev));
  if (!d) lg_fatal("out of memory (grid_dev)");
  d->kind = 0;
  d->prec = HGRID_C128;
  d->n = (size_t)(g->nx * g->ny * g->nz);
  d->valid = GV_HOST;
  g->dev = d;
}

cgrid *cgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL complex (*value_outside)(cgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id) {
  cgrid *grid;
  if (nx < 1 || ny < 1 || nz < 1) lg_fatal("Error in cgrid_alloc(). Invalid dimensions.");
  if (!(grid = (cgrid *)calloc(1, sizeof(cgrid)))) lg_fatal("Error in cgrid_alloc(). Could not allocate memory for grid structure.");
  grid->nx = nx;
  grid->ny = ny;
  grid->nz = nz;
  grid->grid_len = ((size_t)(nx * ny * nz)) * sizeof(REAL complex);
  grid->step = step;
  if (posix_memalign((void **)&grid->value, 256, grid->grid_len))
    lg_fatal("Error in cgrid_alloc(). Could not allocate memory for cgrid->value.");
  grid->x0 = grid->y0 = grid->z0 = 0.0;
  grid->kx0 = grid->ky0 = grid->kz0 = 0.0;
  grid->omega = 0.0;
  grid->fft_norm = 1.0 / (REAL)(grid->nx * grid->ny * grid->nz);
  grid->fft_norm2 = grid->fft_norm;
  if (grid->nx > 1) grid->fft_norm2 *= grid->step;
  if (grid->ny > 1) grid->fft_norm2 *= grid->step;
  if (grid->nz > 1) grid->fft_norm2 *= grid->step;
  strncpy(grid->id, id ? id : "", 32);
  grid->id[31] = 0;
  grid->value_outside = value_outside ? value_outside : cgrid_value_outside_dirichlet;
  if (outside_params_ptr) grid->outside_params_ptr = outside_params_ptr;
  else {
    grid->default_outside_params = 0.0;
    grid->outside_params_ptr = &grid->default_outside_params;
  }
  memset(grid->value, 0, grid->grid_len);
  grid->flag = 0;
  grid->host_lock = 0;
  cdev_init(grid);
  return grid;
}

cgrid *cgrid_clone(cgrid *grid, char *id) {
  cgrid *n = cgrid_alloc(grid->nx, grid->ny, grid->nz, grid->step, grid->value_outside, NULL, id);
  n->outside_params_ptr = grid->outside_params_ptr == &grid->default_outside_params ? &n->default_outside_params
                                                                                     : grid->outside_params_ptr;
  n->default_outside_params = grid->default_outside_params;
  n->x0 = grid->x0; n->y0 = grid->y0; n->z0 = grid->z0;
  n->kx0 = grid->kx0; n->ky0 = grid->ky0; n->kz0 = grid->kz0;
  n->omega = grid->omega;
  n->dev->prec = grid->dev->prec;
  return n;
}

static void free_ktab(ktab_cache *k) {
  for (int a = 0; a < 3; ++a) {
    if (k->t[a]) hgrid_free(k->t[a]);
    if (k->k2[a]) hgrid_free(k->k2[a]);
    if (k->kc[a]) hgrid_free(k->kc[a]);
    k->t[a] = NULL;
    k->k2[a] = NULL;
    k->kc[a] = NULL;
  }
  k->valid = 0;
}

void cgrid_free(cgrid *grid) {
  if (!grid) return;
  grid_dev *d = grid->dev;
  if (d) {
    if (d->d) lg_hip(hgrid_free(d->d), "hipFree");
    if (d->dscal) lg_hip(hgrid_free(d->dscal), "hipFree");
    if (d->work) lg_hip(hgrid_free(d->work), "hipFree");
    if (d->rho) lg_hip(hgrid_free(d->rho), "hipFree");
    if (d->alt) lg_hip(hgrid_free(d->alt), "hipFree");
    free_ktab(&d->kin);
    free_ktab(&d->kin2);
    free_ktab(&d->kin_energy);
    free(d);
  }
  free(grid->value);
  free(grid);
}

void cgrid_claim(cgrid *grid) {
  if (!grid) lg_fatal("Attempting to claim a non-existent grid.");
  if (grid->flag) lg_fatal("Attempting to claim grid twice.");
  grid->flag = 1;
}

void cgrid_release(cgrid *grid) {
  if (!grid->flag) lg_fatal("Attempting to release grid twice.");
  grid->flag = 0;
}

void cgrid_set_origin(cgrid *grid, REAL x0, REAL y0, REAL z0) { grid->x0 = x0; grid->y0 = y0; grid->z0 = z0; }
void cgrid_shift_origin(cgrid *grid, REAL x0, REAL y0, REAL z0) { grid->x0 += x0; grid->y0 += y0; grid->z0 += z0; }
void cgrid_set_momentum(cgrid *grid, REAL kx0, REAL ky0, REAL kz0) { grid->kx0 = kx0; grid->ky0 = ky0; grid->kz0 = kz0; }
void cgrid_set_rotation(cgrid *grid, REAL omega) { grid->omega = omega; }

/* ---- residency ---------------------------------------------------------------------------- */

grid_dev *lg_cdev(cgrid *g) {
  grid_dev *d = g->dev;
  if (g->host_lock) lg_fatal("grid %s is host-locked; this library has no CPU back end", g->id);
  lg_ensure_device();
  if (!d->dscal) {
    lg_hip(hgrid_malloc((void **)&d->dscal, GS_N * sizeof(double)), "hipMalloc (grid scalars)");
    double init[GS_N] = {1.0, 0.0};
    lg_hip(hgrid_memcpy_h2d(d->dscal, init, sizeof(init)), "hipMemcpy (grid scalars)");
    d->pending = 0;
  }
  if (!d->d) lg_hip(hgrid_malloc(&d->d, d->n * lg_prec_size(d->prec)), "hipMalloc (grid)");
  return d;
}

static void upload(cgrid *g) {
  grid_dev *d = g->dev;
  if (d->prec == HGRID_C128) {
    lg_hip(hgrid_memcpy_h2d(d->d, g->value, d->n * 16), "hipMemcpy H2D");
  } else {
    float *tmp = (float *)malloc(d->n * 8);
    if (!tmp) lg_fatal("out of host memory (upload staging)");
    const double *src = (const double *)g->value;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)(2 * d->n); ++i) tmp[i] = (float)src[i];
    lg_hip(hgrid_memcpy_h2d(d->d, tmp, d->n * 8), "hipMemcpy H2D");
    free(tmp);
  }
  d->valid |= GV_DEV;
  d->zpend = 0;
  d->stamp++;
}

static void download(cgrid *g) {
  grid_dev *d = g->dev;
  if (d->prec == HGRID_C128) {
    lg_hip(hgrid_memcpy_d2h(g->value, d->d, d->n * 16), "hipMemcpy D2H");
  } else {
    float *tmp = (float *)malloc(d->n * 8);
    if (!tmp) lg_fatal("out of host memory (download staging)");
    lg_hip(hgrid_memcpy_d2h(tmp, d->d, d->n * 8), "hipMemcpy D2H");
    double *dst = (double *)g->value;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)(2 * d->n); ++i) dst[i] = (double)tmp[i];
    free(tmp);
  }
  d->valid |= GV_HOST;
}

/* Transform a state left along the contiguous axis by a fused step back to real space
 * (unnormalised inverse: the step's kinetic factor already carried 1/N). */
void lg_c_materialize(cgrid *g) {
  grid_dev *d = g->dev;
  if (d->kpend) lg_fatal("internal: a merged relaxation batch left its state between kinetic half-steps");
  if (!d->zpend) return;
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  hgrid_pass p;
  lg_pass_init(&p, &s, d->prec, d->d, d->d, 2);
  p.dir1 = +1;
  lg_run_pass(&p);
  d->zpend = 0;
}

void *lg_c_raw(cgrid *g, const double **scale) {
  grid_dev *d = lg_cdev(g);
  if (!(d->valid & GV_DEV)) upload(g);
  d->stamp++; /* the fused step writes it */
  *scale = d->pending ? d->dscal + GS_SCALE : NULL;
  return d->d;
}

void *lg_c_in(cgrid *g, const double **scale) {
  grid_dev *d = lg_cdev(g);
  if (!(d->valid & GV_DEV)) upload(g);
  lg_c_materialize(g);
  if (scale) *scale = d->pending ? d->dscal + GS_SCALE : NULL;
  else if (d->pending) lg_c_flush(g);
  return d->d;
}

void *lg_c_in_flushed(cgrid *g) { return lg_c_in(g, NULL); }

void *lg_c_inout(cgrid *g, const double **scale) {
  void *p = lg_c_in(g, scale);
  g->dev->stamp++;
  return p;
}

void *lg_c_out(cgrid *g) {
  grid_dev *d = lg_cdev(g);
  d->zpend = 0; /* about to be overwritten */
  d->stamp++;
  return d->d;
}

const void *lg_c_in_as(cgrid *g, int prec) {
  const void *src = lg_c_in_flushed(g);
  grid_dev *d = g->dev;
  /* FP32 storage with either arithmetic: the same bytes */
  if (lg_prec_size(prec) == lg_prec_size(d->prec)) return src;
  if (d->alt && d->alt_prec == prec && d->alt_stamp == d->stamp) return d->alt;
  if (d->alt && lg_prec_size(d->alt_prec) != lg_prec_size(prec)) {
    lg_hip(hgrid_free(d->alt), "hipFree");
    d->alt = NULL;
  }
  if (!d->alt) lg_hip(hgrid_malloc(&d->alt, d->n * lg_prec_size(prec)), "hipMalloc (converted grid)");
  if (prec == HGRID_C128) lg_hip(hgrid_c64_to_c128(d->alt, src, d->n), "convert");
  else lg_hip(hgrid_c128_to_c64(d->alt, src, d->n), "convert");
  d->alt_prec = prec;
  d->alt_stamp = d->stamp;
  return d->alt;
}

/* After a kernel wrote the device array: the host copy is stale.  scale_consumed = the kernel
 * applied the pending scale, so it is reset to 1. */
void lg_c_written(cgrid *g, int scale_consumed) {
  grid_dev *d = g->dev;
  d->valid = GV_DEV;
  d->stamp++;
  d->zpend = 0;
  if (scale_consumed && d->pending) {
    lg_hip(hgrid_scalar_set(d->dscal + GS_SCALE, 1.0, 0.0), "scalar_set");
    d->pending = 0;
  }
}

void lg_c_set_pending(cgrid *g, double re, double im) {
  grid_dev *d = lg_cdev(g);
  if (!(d->valid & GV_DEV)) upload(g);
  lg_hip(hgrid_scalar_mul(d->dscal + GS_SCALE, re, im), "scalar_mul");
  d->pending = 1;
  d->valid = GV_DEV;
  d->stamp++;
}

void lg_c_flush(cgrid *g) {
  grid_dev *d = g->dev;
  if (!d->pending) return;
  lg_hip(hgrid_cscale(d->d, d->prec, (long)d->n, 1.0, 0.0, d->dscal + GS_SCALE), "cscale");
  lg_hip(hgrid_scalar_set(d->dscal + GS_SCALE, 1.0, 0.0), "scalar_set");
  d->pending = 0;
  d->valid = GV_DEV;
}

void lg_c_host_sync(cgrid *g) {
  grid_dev *d = g->dev;
  if (d->valid & GV_HOST) return;
  lg_c_materialize(g);
  lg_c_flush(g);
  download(g);
}

void cgrid_host_sync(cgrid *grid) { lg_c_host_sync(grid); }

void cgrid_host_modified(cgrid *grid) {
  grid_dev *d = grid->dev;
  d->valid = GV_HOST;
  d->stamp++;
  d->zpend = 0;
  if (d->pending) {
    lg_hip(hgrid_scalar_set(d->dscal + GS_SCALE, 1.0, 0.0), "scalar_set");
    d->pending = 0;
  }
}

void cgrid_set_precision(cgrid *grid, char prec) {
  int p = prec == GRID_PREC_SINGLE ? HGRID_C64 : prec == GRID_PREC_SINGLE_STORAGE ? HGRID_C64X : HGRID_C128;
  grid_dev *d = grid->dev;
  if (d->prec == p) return;
  d->stamp++;
  if (lg_prec_size(p) == lg_prec_size(d->prec)) {
    /* FP32 storage either way: only the arithmetic of the passes changes */
    d->prec = p;
    free_ktab(&d->kin);
    free_ktab(&d->kin2);
    return;
  }
  if (d->d) {
    if (d->valid & GV_DEV) {
      lg_c_flush(grid);
      void *nd;
      lg_hip(hgrid_malloc(&nd, d->n * lg_prec_size(p)), "hipMalloc (grid)");
      if (p != HGRID_C128) lg_hip(hgrid_c128_to_c64(nd, d->d, d->n), "convert");
      else lg_hip(hgrid_c64_to_c128(nd, d->d, d->n), "convert");
      lg_hip(hgrid_free(d->d), "hipFree");
      d->d = nd;
    } else {
      lg_hip(hgrid_free(d->d), "hipFree");
      d->d = NULL;
    }
    if (d->work) {
      lg_hip(hgrid_free(d->work), "hipFree");
      d->work = NULL;
    }
    free_ktab(&d->kin);
    free_ktab(&d->kin2);
  }
  d->prec = p;
}

char cgrid_get_precision(cgrid *grid) {
  const int p = grid->dev->prec;
  return p == HGRID_C64 ? GRID_PREC_SINGLE : p == HGRID_C64X ? GRID_PREC_SINGLE_STORAGE : GRID_PREC_DOUBLE;
}

double lg_step_factor(INT nx, INT ny, INT nz, REAL step) {
  double f = 1.0;
  if (nx != 1) f *= step;
  if (ny != 1) f *= step;
  if (nz != 1) f *= step;
  return f;
}

/* ---- host-side access ----------------------------------------------------------------------- */

REAL complex cgrid_value_outside_dirichlet(cgrid *grid, INT i, INT j, INT k) {
  (void)i; (void)j; (void)k;
  return *((REAL complex *)grid->outside_params_ptr);
}

REAL complex cgrid_value_outside_neumann(cgrid *grid, INT i, INT j, INT k) {
  INT nx = grid->nx, ny = grid->ny, nz = grid->nz;
  if (i < 0) i = -i;
  if (i > nx - 1) i = 2 * nx - i - 2;
  if (j < 0) j = -j;
  if (j > ny - 1) j = 2 * ny - j - 2;
  if (k < 0) k = -k;
  if (k > nz - 1) k = 2 * nz - k - 2;
  return cgrid_value_at_index(grid, i, j, k);
}

REAL complex cgrid_value_outside_periodic(cgrid *grid, INT i, INT j, INT k) {
  INT nx = grid->nx, ny = grid->ny, nz = grid->nz;
  i %= nx;
  if (i < 0) i = nx + i;
  j %= ny;
  if (j < 0) j = ny + j;
  k %= nz;
  if (k < 0) k = nz + k;
  return cgrid_value_at_index(grid, i, j, k);
}

REAL complex cgrid_value_at_index(cgrid *grid, INT i, INT j, INT k) {
  if (i < 0 || j < 0 || k < 0 || i >= grid->nx || j >= grid->ny || k >= grid->nz)
    return grid->value_outside(grid, i, j, k);
  lg_c_host_sync(grid);
  return grid->value[(i * grid->ny + j) * grid->nz + k];
}

void cgrid_value_to_index(cgrid *grid, INT i, INT j, INT k, REAL complex value) {
  if (i < 0 || j < 0 || k < 0 || i >= grid->nx || j >= grid->ny || k >= grid->nz)
    lg_fatal("Index out of range in cgrid_value_to_index().");
  lg_c_host_sync(grid);
  grid->value[(i * grid->ny + j) * grid->nz + k] = value;
  grid->dev->valid = GV_HOST;
}

/* ---- device operations -------------------------------------------------------------------- */

void cgrid_zero(cgrid *grid) { cgrid_constant(grid, 0.0); }

void cgrid_constant(cgrid *grid, REAL complex c) {
  grid_dev *d = lg_cdev(grid);
  lg_hip(hgrid_cfill(d->d, d->prec, (long)d->n, creal(c), cimag(c)), "cfill");
  lg_c_written(grid, 1);
}

void cgrid_copy(cgrid *dst, cgrid *src) {
  if (dst->nx != src->nx || dst->ny != src->ny || dst->nz != src->nz) lg_fatal("Incompatible grids in cgrid_copy().");
  /* the grid geometry and the moving frame travel with the values (reference src/cgrid.c:626-632) */
  dst->step = src->step;
  dst->x0 = src->x0;
  dst->y0 = src->y0;
  dst->z0 = src->z0;
  dst->kx0 = src->kx0;
  dst->ky0 = src->ky0;
  dst->kz0 = src->kz0;
  if (dst == src) return;
  void *s = lg_c_in_flushed(src);
  grid_dev *dd = dst->dev;
  if (lg_prec_size(dd->prec) != lg_prec_size(src->dev->prec)) {
    /* copy between FP64 and FP32 storage */
    void *o = lg_c_out(dst);
    if (dd->prec != HGRID_C128) lg_hip(hgrid_c128_to_c64(o, s, dd->n), "convert");
    else lg_hip(hgrid_c64_to_c128(o, s, dd->n), "convert");
  } else {
    void *o = lg_c_out(dst);
    lg_hip(hgrid_memcpy_d2d(o, s, dd->n * lg_prec_size(dd->prec)), "hipMemcpy D2D");
  }
  lg_c_written(dst, 1);
}

void cgrid_multiply(cgrid *grid, REAL complex c) { lg_c_set_pending(grid, creal(c), cimag(c)); }

/* gridc = gridc + d * grida (reference src/cgrid.c:1303-1319) */
void cgrid_add_scaled(cgrid *gridc, REAL complex d, cgrid *grida) {
  if (gridc->nx != grida->nx || gridc->ny != grida->ny || gridc->nz != grida->nz)
    lg_fatal("Incompatible grids in cgrid_add_scaled().");
  if (cgrid_get_precision(grida) != cgrid_get_precision(gridc)) cgrid_set_precision(grida, cgrid_get_precision(gridc));
  const void *a = lg_c_in_flushed(grida);
  void *c = lg_c_in_flushed(gridc);
  lg_hip(hgrid_caxpy(c, a, gridc->dev->prec, (long)gridc->dev->n, creal(d), cimag(d)), "caxpy");
  lg_c_written(gridc, 1);
}

static double fetch_scalar(const double *dptr) {
  double v;
  lg_hip(hgrid_memcpy_d2h(&v, dptr, sizeof(double)), "hipMemcpy D2H (scalar)");
  return v;
}

REAL cgrid_integral_of_square(cgrid *grid) {
  const double *scale;
  void *a = lg_c_in(grid, &scale);
  grid_dev *d = grid->dev;
  long n = (long)d->n;
  double *scr = lg_partials(hgrid_reduce_scratch(n));
  lg_hip(hgrid_sum_norm2(d->dscal + GS_TMP, a, d->prec, n, scale, scr), "sum_norm2");
  return fetch_scalar(d->dscal + GS_TMP) * lg_step_factor(grid->nx, grid->ny, grid->nz, grid->step);
}

REAL complex cgrid_integral_of_conjugate_product(cgrid *grida, cgrid *gridb) {
  void *a = lg_c_in_flushed(grida);
  void *b = lg_c_in_flushed(gridb);
  if (lg_prec_size(grida->dev->prec) != lg_prec_size(gridb->dev->prec))
    lg_fatal("cgrid_integral_of_conjugate_product: grids of different storage precision");
  long n = (long)grida->dev->n;
  double *scr = lg_partials(hgrid_reduce_scratch(n));
  double *out = grida->dev->dscal + GS_TMP;
  lg_hip(hgrid_sum_conj_product(out, a, b, grida->dev->prec, n, scr), "sum_conj_product");
  double v[2];
  lg_hip(hgrid_memcpy_d2h(v, out, 2 * sizeof(double)), "hipMemcpy D2H (scalar)");
  double f = lg_step_factor(grida->nx, grida->ny, grida->nz, grida->step);
  return (v[0] + I * v[1]) * f;
}

/* ---- FFT wrappers (reference src/cgrid.c:1822-1890) --------------------------------------- */

static void check_fft_dims(cgrid *g) {
  INT d[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; ++a)
    if (d[a] > 1 && !hgrid_fft_supported(d[a]))
      lg_fatal("FFT length %ld is not supported (powers of two up to 1024, 3*2^k up to 384)", (long)d[a]);
}

void cgrid_fft(cgrid *grid) {
  check_fft_dims(grid);
  const double *scale;
  void *a = lg_c_inout(grid, &scale);
  lg_shape s;
  lg_squeeze(grid->nx, grid->ny, grid->nz, &s);
  lg_fft_forward(a, a, grid->dev->prec, &s, scale);
  lg_c_written(grid, 1);
}

static void inverse_scaled(cgrid *grid, double post) {
  check_fft_dims(grid);
  const double *scale;
  void *a = lg_c_inout(grid, &scale);
  lg_shape s;
  lg_squeeze(grid->nx, grid->ny, grid->nz, &s);
  lg_fft_inverse(a, a, grid->dev->prec, &s, post, scale);
  lg_c_written(grid, 1);
}

void cgrid_inverse_fft(cgrid *grid) { inverse_scaled(grid, 1.0); }

void cgrid_scaled_inverse_fft(cgrid *grid, REAL complex c) {
  if (cimag(c) == 0.0) inverse_scaled(grid, creal(c));
  else {
    inverse_scaled(grid, 1.0);
    lg_c_set_pending(grid, creal(c), cimag(c));
  }
}

void cgrid_inverse_fft_norm(cgrid *grid) { cgrid_scaled_inverse_fft(grid, grid->fft_norm); }

void cgrid_inverse_fft_norm2(cgrid *grid) { cgrid_scaled_inverse_fft(grid, grid->fft_norm2); }

/* -sum k^2 |psi_k|^2 * fft_norm2, k without the k0 shift (reference src/cgrid_diff.c:817-879) */
REAL cgrid_fft_laplace_expectation_value(cgrid *grid) {
  const double *scale;
  void *a = lg_c_in(grid, &scale);
  ktab_cache *k = lg_k2_tables(grid);
  grid_dev *d = grid->dev;
  lg_hip(lg_sum_k2norm2(d->dscal + GS_TMP, a, d->prec, grid->nx, grid->ny, grid->nz, k->k2[0], k->k2[1], k->k2[2], scale),
         "sum_k2norm2");
  return -fetch_scalar(d->dscal + GS_TMP) * grid->fft_norm2;
}
