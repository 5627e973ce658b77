/*
 * Host-side input and output of grids: functions of the coordinates mapped onto grids, linear
 * interpolation between grids of different sizes, and the reference's binary grid files and
 * ASCII axis cuts.  This is host arithmetic on the host copy, as in the reference, whose GPU
 * build drops the device copy before these loops: writers download first, mappers and readers
 * mark the device copy stale.
 *
 * Reference: src/cgrid.c:360-420 (cgrid_write / cgrid_read), :432-504 (cgrid_write_grid),
 * :793-860 (cgrid_map / cgrid_smooth_map), :2060-2150 (cgrid_value / cgrid_extrapolate);
 * src/rgrid.c:339-398 (rgrid_write / rgrid_read), :505-566 (rgrid_write_grid), :806-870
 * (rgrid_map / rgrid_smooth_map), :2208-2288 (rgrid_value / rgrid_extrapolate); the cell
 * weights of src/cprivate.h:35-68 and src/rprivate.h:65-100.
 */
#include <stdint.h>

#include "internal.h"

/* Coordinate of index i on an axis of n points with origin offset o: (i - n/2) step - o. */
static inline REAL axis_coord(INT i, INT n, REAL step, REAL o) { return ((REAL)(i - n / 2)) * step - o; }

/* ---- cell-averaged sampling ---------------------------------------------------------------
 * With ns samples per axis the value at a point is the average of f over [-step, step]^3 around
 * it, weighted by the product of the hat functions 1 - |s / step|: the samples sit at
 * s_m = -step + (m + 1/2) * (2 step / ns), m < ns (reference src/cprivate.h:35-68).  ns = 1 is
 * the plain point value. */
typedef struct {
  INT ns;
  REAL *off; /* s_m */
  REAL *w;   /* 1 - |s_m / step| */
} cell_rule;

static cell_rule cell_rule_make(REAL step, INT ns) {
  cell_rule r;
  if (ns < 1) lg_fatal("smooth map: ns = %ld must be at least 1", (long)ns);
  r.ns = ns;
  r.off = (REAL *)malloc(sizeof(REAL) * (size_t)ns);
  r.w = (REAL *)malloc(sizeof(REAL) * (size_t)ns);
  if (!r.off || !r.w) lg_fatal("out of host memory (smooth map)");
  const REAL sub = 2.0 * step / (REAL)ns;
  for (INT m = 0; m < ns; ++m) {
    r.off[m] = -step + (((REAL)m) + 0.5) * sub;
    const REAL u = r.off[m] / step;
    r.w[m] = 1.0 - (u < 0.0 ? -u : u);
  }
  return r;
}

static void cell_rule_free(cell_rule *r) {
  free(r->off);
  free(r->w);
}

static REAL complex cell_average_c(REAL complex (*f)(void *, REAL, REAL, REAL), void *arg, REAL x, REAL y, REAL z,
                                   const cell_rule *r) {
  REAL complex acc = 0.0;
  REAL wacc = 0.0;
  for (INT a = 0; a < r->ns; ++a)
    for (INT b = 0; b < r->ns; ++b)
      for (INT c = 0; c < r->ns; ++c) {
        const REAL w = r->w[a] * r->w[b] * r->w[c];
        acc += w * f(arg, x + r->off[a], y + r->off[b], z + r->off[c]);
        wacc += w;
      }
  return acc / wacc;
}

static REAL cell_average_r(REAL (*f)(void *, REAL, REAL, REAL), void *arg, REAL x, REAL y, REAL z, const cell_rule *r) {
  REAL acc = 0.0, wacc = 0.0;
  for (INT a = 0; a < r->ns; ++a)
    for (INT b = 0; b < r->ns; ++b)
      for (INT c = 0; c < r->ns; ++c) {
        const REAL w = r->w[a] * r->w[b] * r->w[c];
        acc += w * f(arg, x + r->off[a], y + r->off[b], z + r->off[c]);
        wacc += w;
      }
  return acc / wacc;
}

/* ---- mapping --------------------------------------------------------------------------------
 * Every point of the host array is written, so nothing is downloaded first; the device copy
 * (and any pending scale or deferred transform) is dropped afterwards. */

static void cfill(cgrid *grid, REAL complex (*func)(void *, REAL, REAL, REAL), void *farg, INT ns) {
  const INT nx = grid->nx, ny = grid->ny, nz = grid->nz;
  const REAL h = grid->step;
  cell_rule r = cell_rule_make(h, ns);
  REAL complex *out = grid->value;
#pragma omp parallel for schedule(static)
  for (INT row = 0; row < nx * ny; ++row) {
    const REAL x = axis_coord(row / ny, nx, h, grid->x0), y = axis_coord(row % ny, ny, h, grid->y0);
    REAL complex *o = out + row * nz;
    for (INT k = 0; k < nz; ++k) {
      const REAL z = axis_coord(k, nz, h, grid->z0);
      o[k] = ns == 1 ? func(farg, x, y, z) : cell_average_c(func, farg, x, y, z, &r);
    }
  }
  cell_rule_free(&r);
  cgrid_host_modified(grid);
}

static void rfill(rgrid *grid, REAL (*func)(void *, REAL, REAL, REAL), void *farg, INT ns) {
  const INT nx = grid->nx, ny = grid->ny, nz = grid->nz, pitch = grid->nz2;
  const REAL h = grid->step;
  cell_rule r = cell_rule_make(h, ns);
  REAL *out = grid->value;
#pragma omp parallel for schedule(static)
  for (INT row = 0; row < nx * ny; ++row) {
    const REAL x = axis_coord(row / ny, nx, h, grid->x0), y = axis_coord(row % ny, ny, h, grid->y0);
    REAL *o = out + row * pitch;
    for (INT k = 0; k < nz; ++k) {
      const REAL z = axis_coord(k, nz, h, grid->z0);
      o[k] = ns == 1 ? func(farg, x, y, z) : cell_average_r(func, farg, x, y, z, &r);
    }
  }
  cell_rule_free(&r);
  rgrid_host_modified(grid);
}

void cgrid_map(cgrid *grid, REAL complex (*func)(void *arg, REAL x, REAL y, REAL z), void *farg) {
  cfill(grid, func, farg, 1);
}

void cgrid_smooth_map(cgrid *grid, REAL complex (*func)(void *arg, REAL x, REAL y, REAL z), void *farg, INT ns) {
  cfill(grid, func, farg, ns);
}

void rgrid_map(rgrid *grid, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg) { rfill(grid, func, farg, 1); }

void rgrid_smooth_map(rgrid *grid, REAL (*func)(void *arg, REAL x, REAL y, REAL z), void *farg, INT ns) {
  rfill(grid, func, farg, ns);
}

/* ---- linear interpolation -----------------------------------------------------------------
 * The cell of coordinate u on an axis: t = (u + origin) / step, i = floor-like truncation of t
 * (one lower for negative t, as the reference does, which for a negative integer t gives the
 * lower cell with fraction 1), fraction t - i, index i + n/2.  Outside indices go through the
 * grid's boundary function. */
static inline void axis_cell(REAL u, REAL origin, REAL step, INT n, INT *i, REAL *frac) {
  REAL t = (u + origin) / step;
  INT c = (INT)t;
  if (t < 0) c--;
  *frac = t - (REAL)c;
  *i = c + n / 2;
}

/* host value at an index of a grid whose host copy is current */
static inline REAL complex c_at(cgrid *g, INT i, INT j, INT k) {
  if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz) return g->value_outside(g, i, j, k);
  return g->value[(i * g->ny + j) * g->nz + k];
}

static inline REAL r_at(rgrid *g, INT i, INT j, INT k) {
  if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz) return g->value_outside(g, i, j, k);
  return g->value[(i * g->ny + j) * g->nz2 + k];
}

static REAL complex c_interp(cgrid *g, REAL x, REAL y, REAL z) {
  INT i, j, k;
  REAL fx, fy, fz;
  axis_cell(x, g->x0, g->step, g->nx, &i, &fx);
  axis_cell(y, g->y0, g->step, g->ny, &j, &fy);
  axis_cell(z, g->z0, g->step, g->nz, &k, &fz);
  const REAL gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
  /* nested form of the trilinear weights (reference src/cgrid.c:2100-2104) */
  return gx * (gy * (gz * c_at(g, i, j, k) + fz * c_at(g, i, j, k + 1)) +
               fy * (gz * c_at(g, i, j + 1, k) + fz * c_at(g, i, j + 1, k + 1))) +
         fx * (gy * (gz * c_at(g, i + 1, j, k) + fz * c_at(g, i + 1, j, k + 1)) +
               fy * (gz * c_at(g, i + 1, j + 1, k) + fz * c_at(g, i + 1, j + 1, k + 1)));
}

static REAL r_interp(rgrid *g, REAL x, REAL y, REAL z) {
  INT i, j, k;
  REAL fx, fy, fz;
  axis_cell(x, g->x0, g->step, g->nx, &i, &fx);
  axis_cell(y, g->y0, g->step, g->ny, &j, &fy);
  axis_cell(z, g->z0, g->step, g->nz, &k, &fz);
  const REAL gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
  /* expanded form of the trilinear weights (reference src/rgrid.c:2254-2255) */
  return gx * gy * gz * r_at(g, i, j, k) + fx * gy * gz * r_at(g, i + 1, j, k) + gx * fy * gz * r_at(g, i, j + 1, k) +
         gx * gy * fz * r_at(g, i, j, k + 1) + fx * fy * gz * r_at(g, i + 1, j + 1, k) +
         fx * gy * fz * r_at(g, i + 1, j, k + 1) + gx * fy * fz * r_at(g, i, j + 1, k + 1) +
         fx * fy * fz * r_at(g, i + 1, j + 1, k + 1);
}

REAL complex cgrid_value(cgrid *grid, REAL x, REAL y, REAL z) {
  lg_c_host_sync(grid);
  return c_interp(grid, x, y, z);
}

REAL rgrid_value(rgrid *grid, REAL x, REAL y, REAL z) {
  rgrid_host_sync(grid);
  return r_interp(grid, x, y, z);
}

void cgrid_extrapolate(cgrid *dest, cgrid *src) {
  if (dest == src) lg_fatal("Source and destination grids cannot be the same in cgrid_extrapolate().");
  lg_c_host_sync(src);
  const INT nx = dest->nx, ny = dest->ny, nz = dest->nz;
  const REAL h = dest->step;
#pragma omp parallel for schedule(static)
  for (INT row = 0; row < nx * ny; ++row) {
    const REAL x = axis_coord(row / ny, nx, h, dest->x0), y = axis_coord(row % ny, ny, h, dest->y0);
    for (INT k = 0; k < nz; ++k) dest->value[row * nz + k] = c_interp(src, x, y, axis_coord(k, nz, h, dest->z0));
  }
  cgrid_host_modified(dest);
}

void rgrid_extrapolate(rgrid *dest, rgrid *src) {
  if (dest == src) lg_fatal("Source and destination grids cannot be the same in rgrid_extrapolate().");
  rgrid_host_sync(src);
  const INT nx = dest->nx, ny = dest->ny, nz = dest->nz, pitch = dest->nz2;
  const REAL h = dest->step;
#pragma omp parallel for schedule(static)
  for (INT row = 0; row < nx * ny; ++row) {
    const REAL x = axis_coord(row / ny, nx, h, dest->x0), y = axis_coord(row % ny, ny, h, dest->y0);
    for (INT k = 0; k < nz; ++k) dest->value[row * pitch + k] = r_interp(src, x, y, axis_coord(k, nz, h, dest->z0));
  }
  rgrid_host_modified(dest);
}

/* ---- binary grid files: [INT nx][INT ny][INT nz][REAL step][values]; complex grids store
 * nx*ny*nz interleaved complex values, real grids their nz2-padded rows ------------------- */

static void write_header(FILE *out, INT nx, INT ny, INT nz, REAL step) {
  const INT dims[3] = {nx, ny, nz};
  if (fwrite(dims, sizeof(INT), 3, out) != 3 || fwrite(&step, sizeof(REAL), 1, out) != 1)
    lg_fatal("short write of a grid header");
}

/* 0 on success */
static int read_header(FILE *in, INT *nx, INT *ny, INT *nz, REAL *step) {
  INT dims[3];
  if (fread(dims, sizeof(INT), 3, in) != 3 || fread(step, sizeof(REAL), 1, in) != 1) return -1;
  *nx = dims[0];
  *ny = dims[1];
  *nz = dims[2];
  if (*nx < 1 || *ny < 1 || *nz < 1) return -1;
  return 0;
}

void cgrid_write(cgrid *grid, FILE *out) {
  lg_c_host_sync(grid);
  write_header(out, grid->nx, grid->ny, grid->nz, grid->step);
  const size_t n = (size_t)(grid->nx * grid->ny * grid->nz);
  if (fwrite(grid->value, sizeof(REAL complex), n, out) != n) lg_fatal("short write in cgrid_write()");
}

void rgrid_write(rgrid *grid, FILE *out) {
  rgrid_host_sync(grid);
  write_header(out, grid->nx, grid->ny, grid->nz, grid->step);
  const size_t n = (size_t)(grid->nx * grid->ny * grid->nz2);
  if (fwrite(grid->value, sizeof(REAL), n, out) != n) lg_fatal("short write in rgrid_write()");
}

cgrid *cgrid_read(cgrid *grid, FILE *in) {
  INT nx, ny, nz;
  REAL step;
  if (read_header(in, &nx, &ny, &nz, &step)) {
    fprintf(stderr, "libgrid: Failed to read grid header in cgrid_read().\n");
    return NULL;
  }
  if (!grid) grid = cgrid_alloc(nx, ny, nz, step, CGRID_PERIODIC_BOUNDARY, NULL, "read_grid");
  const size_t n = (size_t)(nx * ny * nz);
  if (nx == grid->nx && ny == grid->ny && nz == grid->nz && step == grid->step) {
    if (fread(grid->value, sizeof(REAL complex), n, in) != n) lg_fatal("Short read in cgrid_read().");
    cgrid_host_modified(grid);
    return grid;
  }
  /* a different grid on disk: interpolate it onto this one (reference src/cgrid.c:403-416) */
  fprintf(stderr, "libgrid: Grid in file has different size than grid in memory.\n");
  fprintf(stderr, "libgrid: Interpolating between grids.\n");
  cgrid *tmp = cgrid_alloc(nx, ny, nz, step, grid->value_outside, NULL, "cgrid_read_temp");
  if (fread(tmp->value, sizeof(REAL complex), n, in) != n) lg_fatal("Short read in cgrid_read().");
  cgrid_host_modified(tmp);
  cgrid_extrapolate(grid, tmp);
  cgrid_free(tmp);
  return grid;
}

rgrid *rgrid_read(rgrid *grid, FILE *in) {
  INT nx, ny, nz;
  REAL step;
  if (read_header(in, &nx, &ny, &nz, &step)) {
    fprintf(stderr, "libgrid: Failed to read grid header in rgrid_read().\n");
    return NULL;
  }
  if (!grid) grid = rgrid_alloc(nx, ny, nz, step, RGRID_PERIODIC_BOUNDARY, NULL, "read_grid");
  const size_t n = (size_t)(nx * ny * (2 * (nz / 2 + 1)));
  if (nx == grid->nx && ny == grid->ny && nz == grid->nz && step == grid->step) {
    if (fread(grid->value, sizeof(REAL), n, in) != n) lg_fatal("Short read in rgrid_read().");
    rgrid_host_modified(grid);
    return grid;
  }
  fprintf(stderr, "libgrid: Grid in file has different size than grid in memory.\n");
  fprintf(stderr, "libgrid: Interpolating between grids.\n");
  rgrid *tmp = rgrid_alloc(nx, ny, nz, step, grid->value_outside, NULL, "rgrid_read_temp");
  if (fread(tmp->value, sizeof(REAL), n, in) != n) lg_fatal("Short read in rgrid_read().");
  rgrid_host_modified(tmp);
  rgrid_extrapolate(grid, tmp);
  rgrid_free(tmp);
  return grid;
}

/* ---- .grd files with ASCII cuts --------------------------------------------------------------
 * base.grd holds the binary grid; base.x, base.y and base.z hold the cut through the middle
 * point along each axis, one line per point: the coordinate (i - n/2) step (no origin offset),
 * then the value(s), each in "%le". */

static FILE *open_out(const char *base, const char *ext) {
  char file[2048];
  snprintf(file, sizeof(file), "%s.%s", base, ext);
  FILE *fp = fopen(file, "w");
  if (!fp) lg_fatal("Can't open %s for writing.", file);
  return fp;
}

void cgrid_write_grid(char *base, cgrid *grid) {
  lg_c_host_sync(grid);
  FILE *fp = open_out(base, "grd");
  cgrid_write(grid, fp);
  fclose(fp);
  const INT n[3] = {grid->nx, grid->ny, grid->nz};
  static const char *ext[3] = {"x", "y", "z"};
  for (int a = 0; a < 3; ++a) {
    fp = open_out(base, ext[a]);
    INT idx[3] = {n[0] / 2, n[1] / 2, n[2] / 2};
    for (INT m = 0; m < n[a]; ++m) {
      idx[a] = m;
      const REAL complex v = c_at(grid, idx[0], idx[1], idx[2]);
      fprintf(fp, FMT_R " " FMT_R " " FMT_R "\n", ((REAL)(m - n[a] / 2)) * grid->step, creal(v), cimag(v));
    }
    fclose(fp);
  }
}

void rgrid_write_grid(char *base, rgrid *grid) {
  rgrid_host_sync(grid);
  FILE *fp = open_out(base, "grd");
  rgrid_write(grid, fp);
  fclose(fp);
  const INT n[3] = {grid->nx, grid->ny, grid->nz};
  static const char *ext[3] = {"x", "y", "z"};
  for (int a = 0; a < 3; ++a) {
    fp = open_out(base, ext[a]);
    INT idx[3] = {n[0] / 2, n[1] / 2, n[2] / 2};
    for (INT m = 0; m < n[a]; ++m) {
      idx[a] = m;
      fprintf(fp, FMT_R " " FMT_R "\n", ((REAL)(m - n[a] / 2)) * grid->step, r_at(grid, idx[0], idx[1], idx[2]));
    }
    fclose(fp);
  }
}
