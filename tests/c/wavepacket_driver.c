/*
 * A libgrid user program in the shape of the reference's examples/wavepacket.c:111-149: map a
 * Gaussian packet and a harmonic potential onto grids, normalise, then loop over time steps
 * printing the norm, writing |psi|^2 with rgrid_write_grid and propagating with
 * grid_wf_propagate.  It includes <grid/grid.h> as the reference's programs do and uses only the
 * reference's calls; the GPU selection (grid_hip_enable for cuda_enable) is the one change.
 * Here the propagator is WF_2ND_ORDER_FFT, the potential 0.5 k r^2, and the density is written
 * at the first and the last iteration only (the reference writes it every iteration).  At the
 * end the wavefunction goes to <prefix>.final.bin with cgrid_write.
 *
 * usage: wavepacket_driver <threads> <n> <step> <dt> <iters> <kx> <ky> <kz> <w> <xc> <yc> <zc>
 *                          <k> <prefix>
 */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <grid/grid.h>

typedef struct {
  REAL kx, ky, kz, w, xc, yc, zc;
} packet_params;

typedef struct {
  REAL k;
} trap_params;

static REAL complex packet(void *arg, REAL x, REAL y, REAL z) {
  const packet_params *p = (const packet_params *)arg;
  const REAL dx = x - p->xc, dy = y - p->yc, dz = z - p->zc;
  return CEXP(-(dx * dx + dy * dy + dz * dz) / (p->w * p->w) + I * (p->kx * dx + p->ky * dy + p->kz * dz));
}

static REAL complex trap(void *arg, REAL x, REAL y, REAL z) {
  const trap_params *t = (const trap_params *)arg;
  return 0.5 * t->k * (x * x + y * y + z * z);
}

int main(int argc, char *argv[]) {
  if (argc != 15) {
    fprintf(stderr, "usage: wavepacket_driver <threads> <n> <step> <dt> <iters> <kx> <ky> <kz> <w> <xc> <yc> <zc> <k> <prefix>\n");
    return 2;
  }
  const INT threads = atol(argv[1]), n = atol(argv[2]), iters = atol(argv[5]);
  const REAL step = atof(argv[3]), dt = atof(argv[4]);
  packet_params wp = {atof(argv[6]), atof(argv[7]), atof(argv[8]), atof(argv[9]), atof(argv[10]), atof(argv[11]),
                      atof(argv[12])};
  trap_params tp = {atof(argv[13])};
  const char *prefix = argv[14];
  char fname[512];

  grid_threads_init(threads);
  int gpus[1] = {0};
  grid_hip_enable(1, 1, gpus);

  wf *gwf = grid_wf_alloc(n, n, n, step, 1.0, WF_PERIODIC_BOUNDARY, WF_2ND_ORDER_FFT, "WF");
  cgrid *potential = cgrid_alloc(n, n, n, step, CGRID_PERIODIC_BOUNDARY, 0, "potential");
  rgrid *rworkspace = rgrid_alloc(n, n, n, step, RGRID_PERIODIC_BOUNDARY, 0, "rworkspace");
  if (!gwf || !potential || !rworkspace) return 1;

  grid_wf_map(gwf, packet, &wp);
  grid_wf_normalize(gwf);
  cgrid_smooth_map(potential, trap, &tp, 1);

  const REAL complex time = dt;
  for (INT l = 0; l < iters; l++) {
    printf("Iteration " FMT_I " with wf norm = " FMT_R "\n", l, grid_wf_norm(gwf));
    if (l == 0 || l == iters - 1) {
      grid_wf_density(gwf, rworkspace);
      snprintf(fname, sizeof(fname), "%s-" FMT_I, prefix, l);
      rgrid_write_grid(fname, rworkspace);
    }
    grid_wf_propagate(gwf, potential, time);
  }
  printf("Final wf norm = " FMT_R "\n", grid_wf_norm(gwf));

  snprintf(fname, sizeof(fname), "%s.final.bin", prefix);
  FILE *fp = fopen(fname, "wb");
  if (!fp) return 1;
  cgrid_write(gwf->grid, fp);
  fclose(fp);

  grid_wf_free(gwf);
  rgrid_free(rworkspace);
  cgrid_free(potential);
  return 0;
}
