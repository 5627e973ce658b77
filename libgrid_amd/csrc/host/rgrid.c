/*
 * Real grids: the padded host layout of the reference (src/rgrid.c:38-148, nz2 = 2 (nz/2+1))
 * with a dense [nx][ny][nz] FP64 device copy.  Only what the hot path needs: allocation,
 * mapping, element access, a few reductions, and residency.
 */
#include "internal.h"

rgrid *rgrid_alloc(INT nx, INT ny, INT nz, REAL step, REAL (*value_outside)(rgrid *, INT, INT, INT),
                   void *outside_params_ptr, char *id) {
  rgrid *grid;
  if (nx < 1 || ny < 1 || nz < 1) lg_fatal("Error in rgrid_alloc(). Invalid dimensions.");
  if (!(grid = (rgrid *)calloc(1, sizeof(rgrid)))) {
    fprintf(stderr, "libgrid: Error in rgrid_alloc(). Could not allocate memory for grid structure.\n");
    return NULL;
  }
  grid->nx = nx;
  grid->ny = ny;
  grid->nz = nz;
  grid->nz2 = 2 * (nz / 2 + 1);
  grid->grid_len = ((size_t)(nx * ny * grid->nz2)) * sizeof(REAL);
  if (posix_memalign((void **)&grid->value, 256, grid->grid_len)) lg_fatal("Error in rgrid_alloc(). Could not allocate memory for rgrid->value.");
  grid->step = step;
  grid->x0 = grid->y0 = grid->z0 = 0.0;
  grid->kx0 = grid->ky0 = grid->kz0 = 0.0;
  grid->value_outside = value_outside ? value_outside : rgrid_value_outside_dirichlet;
  if (outside_params_ptr) grid->outside_params_ptr = outside_params_ptr;
  else {
    grid->default_outside_params = 0.0;
    grid->outside_params_ptr = &grid->default_outside_params;
  }
  if (grid->value_outside == RGRID_NEUMANN_BOUNDARY)
    grid->fft_norm = 1.0 / (2.0 * ((REAL)grid->nx) * 2.0 * ((REAL)grid->ny) * 2.0 * ((REAL)grid->nz));
  else
    grid->fft_norm = 1.0 / ((REAL)(grid->nx * grid->ny * grid->nz));
  grid->fft_norm2 = grid->fft_norm;
  if (grid->nx > 1) grid->fft_norm2 *= grid->step;
  if (grid->ny > 1) grid->fft_norm2 *= grid->step;
  if (grid->nz > 1) grid->fft_norm2 *= grid->step;
  strncpy(grid->id, id ? id : "", 32);
  grid->id[31] = 0;
  memset(grid->value, 0, grid->grid_len);
  grid->flag = 0;
  grid->host_lock = 0;
  grid_dev *d = (grid_dev *)calloc(1, sizeof(grid_dev));
  if (!d) lg_fatal("out of memory (grid_dev)");
  d->kind = 1;
  d->n = (size_t)(nx * ny * nz);
  d->valid = GV_HOST;
  grid->dev = d;
  return grid;
}

void rgrid_free(rgrid *grid) {
  if (!grid) return;
  grid_dev *d = grid->dev;
  if (d) {
    if (d->d) lg_hip(hgrid_free(d->d), "hipFree");
    if (d->dscal) lg_hip(hgrid_free(d->dscal), "hipFree");
    free(d);
  }
  free(grid->value);
  free(grid);
}

grid_dev *lg_rdev(rgrid *g) {
  grid_dev *d = g->dev;
  if (g->host_lock) lg_fatal("grid %s is host-locked; this library has no CPU back end", g->id);
  lg_ensure_device();
  if (!d->d) lg_hip(hgrid_malloc(&d->d, d->n * sizeof(double)), "hipMalloc (rgrid)");
  if (!d->dscal) lg_hip(hgrid_malloc((void **)&d->dscal, GS_N * sizeof(double)), "hipMalloc (rgrid scalars)");
  return d;
}

/* Upload packs the nz2-padded host rows through a device staging buffer of the padded size. */
static void rupload(rgrid *g) {
  grid_dev *d = g->dev;
  size_t padded = (size_t)(g->nx * g->ny * g->nz2);
  if (g->nz2 == g->nz) {
    lg_hip(hgrid_memcpy_h2d(d->d, g->value, padded * sizeof(double)), "hipMemcpy H2D");
  } else {
    double *stage;
    lg_hip(hgrid_malloc((void **)&stage, padded * sizeof(double)), "hipMalloc (staging)");
    lg_hip(hgrid_memcpy_h2d(stage, g->value, padded * sizeof(double)), "hipMemcpy H2D");
    lg_hip(hgrid_rpack((double *)d->d, stage, g->nx * g->ny, g->nz, g->nz2), "rpack");
    lg_hip(hgrid_free(stage), "hipFree");
  }
  d->valid |= GV_DEV;
}

static void rdownload(rgrid *g) {
  grid_dev *d = g->dev;
  size_t padded = (size_t)(g->nx * g->ny * g->nz2);
  if (g->nz2 == g->nz) {
    lg_hip(hgrid_memcpy_d2h(g->value, d->d, padded * sizeof(double)), "hipMemcpy D2H");
  } else {
    double *stage;
    lg_hip(hgrid_malloc((void **)&stage, padded * sizeof(double)), "hipMalloc (staging)");
    lg_hip(hgrid_runpack(stage, (const double *)d->d, g->nx * g->ny, g->nz, g->nz2), "runpack");
    lg_hip(hgrid_memcpy_d2h(g->value, stage, padded * sizeof(double)), "hipMemcpy D2H");
    lg_hip(hgrid_free(stage), "hipFree");
  }
  d->valid |= GV_HOST;
}

const double *lg_r_in(rgrid *g) {
  grid_dev *d = lg_rdev(g);
  if (!(d->valid & GV_DEV)) rupload(g);
  return (const double *)d->d;
}

double *lg_r_out(rgrid *g) { return (double *)lg_rdev(g)->d; }

void lg_r_written(rgrid *g) { g->dev->valid = GV_DEV; }

void rgrid_host_sync(rgrid *grid) {
  if (grid->dev->valid & GV_HOST) return;
  rdownload(grid);
}

void rgrid_host_modified(rgrid *grid) { grid->dev->valid = GV_HOST; }

REAL rgrid_value_outside_dirichlet(rgrid *grid, INT i, INT j, INT k) {
  (void)i; (void)j; (void)k;
  return *((REAL *)grid->outside_params_ptr);
}

REAL rgrid_value_outside_neumann(rgrid *grid, INT i, INT j, INT k) {
  INT nx = grid->nx, ny = grid->ny, nz = grid->nz;
  if (i < 0) i = -i;
  if (i > nx - 1) i = 2 * nx - i - 2;
  if (j < 0) j = -j;
  if (j > ny - 1) j = 2 * ny - j - 2;
  if (k < 0) k = -k;
  if (k > nz - 1) k = 2 * nz - k - 2;
  return rgrid_value_at_index(grid, i, j, k);
}

REAL rgrid_value_outside_periodic(rgrid *grid, INT i, INT j, INT k) {
  INT nx = grid->nx, ny = grid->ny, nz = grid->nz;
  i %= nx;
  if (i < 0) i = nx + i;
  j %= ny;
  if (j < 0) j = ny + j;
  k %= nz;
  if (k < 0) k = nz + k;
  return rgrid_value_at_index(grid, i, j, k);
}

REAL rgrid_value_at_index(rgrid *grid, INT i, INT j, INT k) {
  if (i < 0 || j < 0 || k < 0 || i >= grid->nx || j >= grid->ny || k >= grid->nz)
    return grid->value_outside(grid, i, j, k);
  rgrid_host_sync(grid);
  return grid->value[(i * grid->ny + j) * grid->nz2 + k];
}

void rgrid_value_to_index(rgrid *grid, INT i, INT j, INT k, REAL value) {
  if (i < 0 || j < 0 || k < 0 || i >= grid->nx || j >= grid->ny || k >= grid->nz)
    lg_fatal("Index out of range in rgrid_value_to_index().");
  rgrid_host_sync(grid);
  grid->value[(i * grid->ny + j) * grid->nz2 + k] = value;
  grid->dev->valid = GV_HOST;
}

void rgrid_zero(rgrid *grid) { rgrid_constant(grid, 0.0); }

void rgrid_constant(rgrid *grid, REAL c) {
  /* host fill keeps the padding at zero as the reference's allocation does */
  INT nxy = grid->nx * grid->ny, nz = grid->nz, nzz = grid->nz2;
#pragma omp parallel for schedule(static)
  for (INT ij = 0; ij < nxy; ij++)
    for (INT k = 0; k < nz; k++) grid->value[ij * nzz + k] = c;
  grid->dev->valid = GV_HOST;
}

static double rsum(rgrid *grid, int squared) {
  const double *a = lg_r_in(grid);
  long n = (long)grid->dev->n;
  double *scr = lg_partials(hgrid_reduce_scratch(n));
  lg_hip(hgrid_sum_real(grid->dev->dscal, a, n, squared, scr), "sum_real");
  double v;
  lg_hip(hgrid_memcpy_d2h(&v, grid->dev->dscal, sizeof(double)), "hipMemcpy D2H (scalar)");
  return v;
}

REAL rgrid_integral(rgrid *grid) { return rsum(grid, 0) * lg_step_factor(grid->nx, grid->ny, grid->nz, grid->step); }

REAL rgrid_integral_of_square(rgrid *grid) {
  return rsum(grid, 1) * lg_step_factor(grid->nx, grid->ny, grid->nz, grid->step);
}
