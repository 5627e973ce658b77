/*
 * Wavefunctions: allocation, the 2nd-order FFT split-operator step, the potential and
 * kinetic phases, norm, energies and the fused Gross-Pitaevskii step.
 *
 * grid_wf_propagate (reference src/grid_wf.c:379-432, branch WF_2ND_ORDER_FFT without an
 * absorbing boundary) computes
 *     psi <- IFFT(K FFT(psi)) / N ; psi <- psi exp(-i t (cons + V) / hbar) ; psi <- IFFT(K FFT(psi)) / N
 * with K = exp(-i (t/2) hbar k^2 / 2m) and cons = -(hbar^2 / 2m) |k0|^2, in that order.  For
 * a grid with at least two non-trivial axes it runs as 2 * m + 3 pencil passes (9 for 3-D):
 * the kinetic factor sits between a forward and an inverse transform of the outermost axis,
 * the potential factor between an inverse and a forward transform of the contiguous axis,
 * and the 1/N of both normalised inverse transforms is folded into the kinetic tables.
 */
#include "internal.h"
#include "relax_rule.h"

char grid_analyze_method = (char)-1; /* 0 = finite difference, 1 = FFT, -1 = not set */

void grid_wf_analyze_method(char method) { grid_analyze_method = method; }

wf *grid_wf_alloc(INT nx, INT ny, INT nz, REAL step, REAL mass, char boundary, char propagator, char *id) {
  wf *gwf;
  REAL complex (*value_outside)(struct cgrid_struct *grid, INT i, INT j, INT k) = NULL;
  if (boundary < WF_DIRICHLET_BOUNDARY || boundary > WF_PERIODIC_BOUNDARY) {
    fprintf(stderr, "libgrid: Error in grid_wf_alloc(). Unknown boundary condition.\n");
    return NULL;
  }
  if (propagator < WF_2ND_ORDER_FFT || propagator > WF_1ST_ORDER_EULER) {
    fprintf(stderr, "libgrid: Error in grid_wf_alloc(). Unknown propagator.\n");
    return NULL;
  }
  if (boundary == WF_DIRICHLET_BOUNDARY && propagator < WF_2ND_ORDER_CN) {
    fprintf(stderr, "libgrid: Error in grid_wf_alloc(). Invalid boundary condition - propagator combination. Dirichlet "
                    "condition can only be used with Crank-Nicolson propagator.\n");
    return NULL;
  }
  if (!(gwf = (wf *)calloc(1, sizeof(wf)))) {
    fprintf(stderr, "libgrid: Error in grid_wf_alloc(). Could not allocate memory for wf.\n");
    return NULL;
  }
  switch (boundary) {
    case WF_DIRICHLET_BOUNDARY: value_outside = CGRID_DIRICHLET_BOUNDARY; break;
    case WF_NEUMANN_BOUNDARY: value_outside = CGRID_NEUMANN_BOUNDARY; break;
    case WF_PERIODIC_BOUNDARY: value_outside = CGRID_PERIODIC_BOUNDARY; break;
  }
  gwf->grid = cgrid_alloc(nx, ny, nz, step, value_outside, 0, id);
  if (grid_analyze_method == -1) grid_analyze_method = 0; /* reference src/cgrid.c:141-149 */
  gwf->mass = mass;
  gwf->norm = 1.0;
  gwf->boundary = boundary;
  gwf->propagator = propagator;
  gwf->cworkspace = gwf->cworkspace2 = gwf->cworkspace3 = NULL;
  gwf->ts_func = NULL;
  gwf->cfft_width = 1.0;
  gwf->kmax = 9999.0;
  return gwf;
}

wf *grid_wf_clone(wf *gwf, char *id) {
  wf *nwf = (wf *)malloc(sizeof(wf));
  if (!nwf) lg_fatal("Out of memory in grid_wf_clone().");
  memcpy(nwf, gwf, sizeof(wf));
  nwf->grid = cgrid_clone(gwf->grid, id);
  nwf->cworkspace = nwf->cworkspace2 = nwf->cworkspace3 = NULL;
  return nwf;
}

void grid_wf_free(wf *gwf) {
  if (!gwf) return;
  if (gwf->grid) cgrid_free(gwf->grid);
  if (gwf->cworkspace) cgrid_free(gwf->cworkspace);
  if (gwf->cworkspace2) cgrid_free(gwf->cworkspace2);
  if (gwf->cworkspace3) cgrid_free(gwf->cworkspace3);
  free(gwf);
}

/* ---- absorbing boundary (reference src/grid_wf.c:99-120, :185-203, :728-755) --------------- */

REAL grid_wf_absorb(INT i, INT j, INT k, INT lx, INT hx, INT ly, INT hy, INT lz, INT hz) {
  REAL t = 0.0;
  if (i < lx) t += ((REAL)(lx - i)) / (REAL)lx;
  else if (i > hx) t += ((REAL)(i - hx)) / (REAL)lx; /* the reference divides by lx on both sides */
  if (j < ly) t += ((REAL)(ly - j)) / (REAL)ly;
  else if (j > hy) t += ((REAL)(j - hy)) / (REAL)ly;
  if (k < lz) t += ((REAL)(lz - k)) / (REAL)lz;
  else if (k > hz) t += ((REAL)(k - hz)) / (REAL)lz;
  t *= 2.0 / 3.0; /* mean of the three directions, doubled: full imaginary time half-way in */
  return t > 1.0 ? 1.0 : t;
}

void grid_wf_boundary(wf *gwf, wf *gwfp, INT lx, INT hx, INT ly, INT hy, INT lz, INT hz) {
  if (!lx && !hx && !ly && !hy && !lz && !hz) {
    gwf->ts_func = NULL;
    gwf->lx = gwf->hx = gwf->ly = gwf->hy = gwf->lz = gwf->hz = 0;
    if (gwfp) {
      gwfp->ts_func = NULL;
      gwfp->lx = 0; /* as the reference: only lx of the predict wavefunction is cleared */
    }
    return;
  }
  gwf->lx = lx;
  gwf->hx = hx;
  gwf->ly = ly;
  gwf->hy = hy;
  gwf->lz = lz;
  gwf->hz = hz;
  gwf->ts_func = &grid_wf_absorb;
  if (gwfp) gwfp->ts_func = &grid_wf_absorb;
}

void grid_wf_merge(wf *dst, cgrid *wfr, cgrid *wfi) {
  cgrid *g = dst->grid;
  if (wfr->nx != g->nx || wfr->ny != g->ny || wfr->nz != g->nz || wfi->nx != g->nx || wfi->ny != g->ny ||
      wfi->nz != g->nz)
    lg_fatal("grid_wf_merge: grids differ in size");
  const int prec = cgrid_get_precision(g);
  if (cgrid_get_precision(wfr) != prec) cgrid_set_precision(wfr, prec);
  if (cgrid_get_precision(wfi) != prec) cgrid_set_precision(wfi, prec);
  const void *r = lg_c_in_flushed(wfr);
  const void *im = lg_c_in_flushed(wfi);
  if (g != wfr) lg_c_in_flushed(g); /* brings a pending z transform / scale to a known state */
  void *out = lg_c_out(g);
  const long bounds[6] = {dst->lx, dst->hx, dst->ly, dst->hy, dst->lz, dst->hz};
  lg_hip(hgrid_wf_merge(out, r, im, g->dev->prec, g->nx, g->ny, g->nz, bounds), "wf_merge");
  lg_c_written(g, 1);
}

void grid_wf_zero(wf *gwf) { cgrid_zero(gwf->grid); }
void grid_wf_constant(wf *gwf, REAL complex c) { cgrid_constant(gwf->grid, c); }
void grid_wf_map(wf *gwf, REAL complex (*func)(void *arg, REAL x, REAL y, REAL z), void *farg) {
  cgrid_map(gwf->grid, func, farg);
}

REAL grid_wf_norm(wf *gwf) { return cgrid_integral_of_square(gwf->grid); }

REAL grid_wf_normalize(wf *gwf) {
  REAL norm = grid_wf_norm(gwf);
  cgrid_multiply(gwf->grid, SQRT(gwf->norm / norm));
  return norm;
}

REAL complex grid_wf_overlap(wf *gwfa, wf *gwfb) { return cgrid_integral_of_conjugate_product(gwfa->grid, gwfb->grid); }

/* ---- pointwise operators ------------------------------------------------------------------ */

void grid_wf_propagate_potential(wf *gwf, REAL complex tstep, cgrid *potential, REAL cons) {
  if (!potential) return;
  cgrid *g = gwf->grid;
  if (potential->nx != g->nx || potential->ny != g->ny || potential->nz != g->nz)
    lg_fatal("grid_wf_propagate_potential: potential and wavefunction grids differ in size");
  /* the FFT propagators ignore ts_func here (reference src/grid_wf.c:584); the position-
   * dependent time of the CN propagators is not built */
  if (gwf->ts_func && gwf->propagator >= WF_2ND_ORDER_CN)
    lg_fatal("grid_wf_propagate_potential: absorbing boundaries with the CN propagators are not supported by this build");
  /* the potential keeps its own precision; a copy in the state's storage precision is used */
  const void *pot = lg_c_in_as(potential, g->dev->prec);
  const double *scale;
  void *psi = lg_c_inout(g, &scale);
  /* c = -i tstep / hbar */
  const double c_re = cimag(tstep) / HBAR, c_im = -creal(tstep) / HBAR;
  lg_hip(hgrid_potential_phase(psi, pot, g->dev->prec, (long)g->dev->n, c_re, c_im, cons, scale), "potential_phase");
  lg_c_written(g, 1);
}

void grid_wf_propagate_kinetic_fft(wf *gwf, REAL complex time) {
  cgrid *g = gwf->grid;
  const REAL complex tm = -I * time * HBAR / (gwf->mass * 2.0); /* reference src/grid_wf_fft.c:146 */
  const double *scale;
  void *psi = lg_c_inout(g, &scale);
  ktab_cache *k = lg_kinetic_tables(g, creal(tm), cimag(tm), 1.0);
  lg_hip(hgrid_kinetic_phase(psi, g->dev->prec, g->nx, g->ny, g->nz, k->t[0], k->t[1], k->t[2], scale, NULL, NULL,
                             NULL, 0.0),
         "kinetic_phase");
  lg_c_written(g, 1);
}

/* reference src/grid_wf_fft.c:206-262: the kinetic phase with components |k - k0| >= kmax
 * set to zero (the CFFT propagator's high-wavenumber sink) */
void grid_wf_propagate_kinetic_cfft(wf *gwf, REAL complex time) {
  cgrid *g = gwf->grid;
  const REAL complex tm = -I * time * HBAR / (gwf->mass * 2.0);
  const double *scale;
  void *psi = lg_c_inout(g, &scale);
  ktab_cache *k = lg_kinetic_tables(g, creal(tm), cimag(tm), 1.0);
  lg_hip(hgrid_kinetic_phase(psi, g->dev->prec, g->nx, g->ny, g->nz, k->t[0], k->t[1], k->t[2], scale, k->kc[0],
                             k->kc[1], k->kc[2], gwf->kmax * gwf->kmax),
         "kinetic_phase (cutoff)");
  lg_c_written(g, 1);
}

void grid_wf_density(wf *gwf, rgrid *density) {
  cgrid *g = gwf->grid;
  if (density->nx != g->nx || density->ny != g->ny || density->nz != g->nz) lg_fatal("grid_wf_density: size mismatch");
  const double *scale;
  const void *psi = lg_c_in(g, &scale);
  double *out = lg_r_out(density);
  lg_hip(hgrid_density(out, psi, g->dev->prec, (long)g->dev->n, scale), "density");
  lg_r_written(density);
}

void grid_real_to_complex_re(cgrid *dest, rgrid *source) {
  if (dest->nx != source->nx || dest->ny != source->ny || dest->nz != source->nz)
    lg_fatal("grid_real_to_complex_re: size mismatch");
  dest->step = source->step;
  const double *src = lg_r_in(source);
  void *dst = lg_c_out(dest);
  lg_hip(hgrid_real_to_complex(dst, dest->dev->prec, src, (long)dest->dev->n), "real_to_complex");
  lg_c_written(dest, 1);
}

static double fetch(const double *dptr) {
  double v;
  lg_hip(hgrid_memcpy_d2h(&v, dptr, sizeof(double)), "hipMemcpy D2H (scalar)");
  return v;
}

REAL grid_grid_expectation_value(cgrid *dgrid, rgrid *opgrid) {
  if (dgrid->nx != opgrid->nx || dgrid->ny != opgrid->ny || dgrid->nz != opgrid->nz)
    lg_fatal("grid_grid_expectation_value: size mismatch");
  const double *scale;
  const void *psi = lg_c_in(dgrid, &scale);
  const double *v = lg_r_in(opgrid);
  grid_dev *d = dgrid->dev;
  long n = (long)d->n;
  double *scr = lg_partials(hgrid_reduce_scratch(n));
  lg_hip(hgrid_sum_vnorm2(d->dscal + GS_TMP, psi, d->prec, v, n, scale, scr), "sum_vnorm2");
  return fetch(d->dscal + GS_TMP) * lg_step_factor(dgrid->nx, dgrid->ny, dgrid->nz, dgrid->step);
}

/* ---- pass plumbing ------------------------------------------------------------------------ */

/* table pointers of squeezed axis a (trivial / padding axes use a trivial physical axis,
 * whose table holds the single value 1 or 0) */
static int trivial_phys(const cgrid *g) { return g->nx == 1 ? 0 : g->ny == 1 ? 1 : 2; }

static const void *ktab_sq(const ktab_cache *k, const lg_shape *s, const cgrid *g, int a) {
  return k->t[s->phys[a] >= 0 ? s->phys[a] : trivial_phys(g)];
}
static const double *k2_sq(const ktab_cache *k, const lg_shape *s, const cgrid *g, int a) {
  return k->k2[s->phys[a] >= 0 ? s->phys[a] : trivial_phys(g)];
}

/* the two other squeezed axes (c1, c2) of a pass on `axis`, as the kernels define them */
static void other_axes(int axis, int *a1, int *a2) {
  if (axis == 0) { *a1 = 1; *a2 = 2; }
  else if (axis == 1) { *a1 = 0; *a2 = 2; }
  else { *a1 = 0; *a2 = 1; }
}

static void set_kin(hgrid_pass *p, const ktab_cache *k, const lg_shape *s, const cgrid *g) {
  int a1, a2;
  other_axes(p->axis, &a1, &a2);
  p->ktab_a = ktab_sq(k, s, g, p->axis);
  p->ktab_1 = ktab_sq(k, s, g, a1);
  p->ktab_2 = ktab_sq(k, s, g, a2);
}

static void set_k2(hgrid_pass *p, const ktab_cache *k, const lg_shape *s, const cgrid *g) {
  int a1, a2;
  other_axes(p->axis, &a1, &a2);
  p->k2_a = k2_sq(k, s, g, p->axis);
  p->k2_1 = k2_sq(k, s, g, a1);
  p->k2_2 = k2_sq(k, s, g, a2);
  /* the pass axis's entries are formed in the kernel: the same expression as lg_k2_tables */
  const long L = p->axis == 0 ? p->n0 : p->axis == 1 ? p->n1 : p->n2;
  p->k2_dk = 2.0 * M_PI / (g->step * (double)L);
}

static void set_kc(hgrid_pass *p, const ktab_cache *k, const lg_shape *s, const cgrid *g, double kcut2) {
  int a1, a2;
  other_axes(p->axis, &a1, &a2);
  p->flags |= HGRID_PF_KCUT;
  p->kc_a = k->kc[s->phys[p->axis] >= 0 ? s->phys[p->axis] : trivial_phys(g)];
  p->kc_1 = k->kc[s->phys[a1] >= 0 ? s->phys[a1] : trivial_phys(g)];
  p->kc_2 = k->kc[s->phys[a2] >= 0 ? s->phys[a2] : trivial_phys(g)];
  p->kcut2 = kcut2;
}

static void *wf_work(cgrid *g) {
  grid_dev *d = g->dev;
  if (!d->work) lg_hip(hgrid_malloc(&d->work, d->n * lg_prec_size(d->prec)), "hipMalloc (wf workspace)");
  return d->work;
}

/* ---- fused step --------------------------------------------------------------------------- */

typedef struct {
  int gpe;               /* 0 complex potential (or none), 1 vext + g rho */
  const void *pot;       /* complex potential on the device (NULL = no potential) */
  const double *vext;
  double g;
  double c_re, c_im, cons;
  double tm_re, tm_im;   /* time_mass of the half step */
  int no_defer;          /* leave the result in real space (no z transform kept pending) */
  const void *pot6;      /* 4th order: exp(c6 (cons + pot6)) applied before the first and after
                          * the last transform of the step (its outer V(t/6) stages) */
  double c6_re, c6_im;
  int kcut;              /* CFFT: zero the components with |k - k0|^2 >= kcut2 */
  double kcut2;
  int want_energy;       /* reductions for E of the input state */
  int want_post_norm;    /* sum |psi|^2 of the output */
  double renorm_target;  /* > 0: make the output's pending scale renormalise it to this norm */
  /* the split stages of a merged relaxation batch (3-D deferred steps, linear potential):
   * 0 a whole step K(t/2) V(t) K(t/2);
   * 1 its start K(t/2) V(t), from psi;
   * 2 K(t) V(t): the last half-kinetic factor of one step and the first of the next, merged;
   * 3 the final K(t/2), back to psi.  Stages 1 and 2 leave the state between kinetic half-steps
   *   (kpend), which stage 3 resolves before anything else may read the grid. */
  int stage;
} step_args;

static void fused_step(wf *gwf, const step_args *sa) {
  cgrid *g = gwf->grid;
  grid_dev *d = g->dev;
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  const double *scale;
  void *psi = lg_c_raw(g, &scale);
  const int prec = d->prec, aK = 3 - s.m;
  /* 3-D: the state stays transformed along the contiguous axis between steps (zpend); the
   * step's first forward and last inverse transform along that axis cancel, and a read-only
   * pass recovers the real-space quantities of the start-of-step state instead (skipped when
   * nothing needs them).  With FP32 butterflies the two roundings fewer per step also lower
   * the error against FP64 (512^3 vortex ring after 500 steps: 1.3e-4 instead of 1.6e-4). */
  const int defer = s.m == 3 && !sa->no_defer;
  const int stage = sa->stage;
  if (stage && (!defer || sa->gpe != 1 || sa->g != 0.0 || (stage >= 2 && (!d->zpend || !d->kpend || sa->want_energy))))
    lg_fatal("internal: merged relaxation stage %d on a state it does not apply to", stage);
  if (stage <= 1 && d->kpend) lg_fatal("internal: a merged relaxation batch was left unfinished");
  if (d->zpend && !defer) lg_c_materialize(g);
  const int zin = d->zpend;
  /* stage 2 applies the kinetic factor of a whole step: the tables of twice the half-step time
   * (one unnormalised round trip per step, so one 1/N) */
  ktab_cache *kt = stage == 2 ? lg_kinetic_tables_in(g, &d->kin2, 2.0 * sa->tm_re, 2.0 * sa->tm_im, g->fft_norm)
                              : lg_kinetic_tables(g, sa->tm_re, sa->tm_im, g->fft_norm);
  ktab_cache *k2 = (sa->want_energy) ? lg_k2_tables(g) : NULL;
  /* V_eff = vext + g |psi|^2 is materialised only for g != 0; with g = 0 the potential pass
   * reads vext itself */
  const int veff = sa->gpe && sa->g != 0.0;
  if (veff && !d->rho) lg_hip(hgrid_malloc((void **)&d->rho, d->n * sizeof(double)), "hipMalloc (density)");

  const long nb2 = hgrid_pass_blocks(prec, s.n[0], s.n[1], s.n[2], 2);
  const long nbK = hgrid_pass_blocks(prec, s.n[0], s.n[1], s.n[2], aK);
  const long nb1 = s.m == 3 ? hgrid_pass_blocks(prec, s.n[0], s.n[1], s.n[2], 1) : 0;
  const long nb9 = nb1 > nb2 ? nb1 : nb2; /* the output norm comes from the last y pass when deferred */
  double *part = lg_partials(HGRID_RED_NCHAN * (nb2 + nbK + nb9));
  double *part1 = part, *partK = part + HGRID_RED_NCHAN * nb2, *part9 = partK + HGRID_RED_NCHAN * nbK;
  hgrid_pass p;

  /* P1: forward along the contiguous axis; pending scale, density and input reductions.  With
   * the state already transformed (zin) the pass instead inverse-transforms in registers for
   * the density / reductions and stores nothing; the first y pass then applies the scale. */
  lg_pass_init(&p, &s, prec, psi, psi, 2);
  if (scale) { p.flags |= HGRID_PF_PRESCALE; p.scale = scale; }
  if (veff) {
    /* V_eff = vext + g |psi|^2 of the start-of-step state, read back by the potential pass */
    p.flags |= HGRID_PF_PRE_VEFF;
    p.rho_out = d->rho;
    p.vext = sa->vext;
    p.g = sa->g;
  }
  if (sa->want_energy) { p.flags |= HGRID_PF_PRE_RED; p.vext = sa->vext; p.partials = part1; }
  if (sa->pot6) {
    /* V(t/6), then the forward transform (never with zin: the 4th-order step is not deferred) */
    p.flags |= HGRID_PF_MID_POT | HGRID_PF_POT_NOFFT1;
    p.dir2 = -1;
    p.pot = sa->pot6;
    p.c_re = sa->c6_re;
    p.c_im = sa->c6_im;
    p.cons = sa->cons;
    p.mid = 1.0;
  }
  if (stage >= 2) {
    /* no start-of-step pass: the stored transform S was produced by a forward transform along
     * the contiguous axis (the previous stage's potential pass), so the first y pass reads it
     * as it is and applies the pending scale */
  } else if (zin) {
    p.dir1 = +1;
    p.flags |= HGRID_PF_PRE_AFTER | HGRID_PF_NOSTORE;
    if (veff || sa->want_energy) lg_run_pass(&p); /* otherwise nothing reads the input state */
  } else {
    lg_run_pass(&p);
  }
  /* with zin the first y pass reads the stored transform S directly.  The forward transform
   * of the real-space state IDFT(S) it replaces would have given n2 * S (an unnormalised round
   * trip), which the kinetic factors' 1/N^2 accounts for: that pass applies n2 and the scale */
  const double *yscale = zin ? scale : NULL;
  int ypost = zin && stage < 2;
  /* all reductions of the step and the pending-scale update go into one finalize launch */
  hgrid_finalize fin;
  memset(&fin, 0, sizeof(fin));
  if (sa->want_energy) {
    hgrid_reduce_job j = {part1, nb2, HGRID_RED_NORM, 3, d->dscal + GS_RED};
    fin.job[fin.njobs++] = j;
  }

  for (int half = 0; half < 2; ++half) {
    if ((stage == 1 || stage == 2) && half == 1) break; /* the step ends after its potential */
    if (stage == 3 && half == 0) continue;               /* only the final kinetic half */
    for (int axis = 1; axis > aK; --axis) {
      lg_pass_init(&p, &s, prec, psi, psi, axis);
      if (yscale) {
        p.flags |= HGRID_PF_PRESCALE;
        p.scale = yscale;
        yscale = NULL;
      }
      if (ypost) {
        p.post = (double)s.n[2];
        ypost = 0;
      }
      lg_run_pass(&p);
    }
    /* kinetic: forward, K / N, inverse along the outermost axis */
    lg_pass_init(&p, &s, prec, psi, psi, aK);
    p.nfft = 2;
    p.dir1 = -1;
    p.dir2 = +1;
    p.flags |= HGRID_PF_MID_KIN;
    if (sa->tm_im == 0.0) p.flags |= HGRID_PF_KIN_REAL; /* exp(tm k^2) with real tm: real tables */
    set_kin(&p, kt, &s, g);
    if (sa->kcut) set_kc(&p, kt, &s, g, sa->kcut2);
    if (half == 0 && sa->want_energy) {
      p.flags |= HGRID_PF_MID_KRED;
      set_k2(&p, k2, &s, g);
      p.partials = partK;
    }
    lg_run_pass(&p);
    if (half == 0 && sa->want_energy) {
      hgrid_reduce_job j = {partK, nbK, HGRID_RED_KIN, 1, d->dscal + GS_RED + HGRID_RED_KIN};
      fin.job[fin.njobs++] = j;
    }
    for (int axis = aK + 1; axis <= 1; ++axis) {
      lg_pass_init(&p, &s, prec, psi, psi, axis);
      p.dir1 = +1;
      if (defer && half == 1 && axis == 1 && sa->want_post_norm) {
        /* sum |S|^2 of the output S, still transformed along the contiguous axis: the state
         * is IDFT(S), whose sum is n2 times larger (Parseval); the finalize multiplies it in */
        p.flags |= HGRID_PF_POST_RED;
        p.partials = part9;
      }
      lg_run_pass(&p);
    }
    if (half == 0) {
      /* potential: inverse, exp(c (cons + V)), forward along the contiguous axis */
      if (sa->gpe || sa->pot) {
        lg_pass_init(&p, &s, prec, psi, psi, 2);
        p.nfft = 2;
        p.dir1 = +1;
        p.dir2 = -1;
        p.c_re = sa->c_re;
        p.c_im = sa->c_im;
        p.cons = sa->cons;
        p.mid = 1.0;
        if (veff) {
          p.flags |= HGRID_PF_MID_GPE;
          p.vext = NULL; /* d->rho holds V_eff */
          p.rho = d->rho;
        } else if (sa->gpe) {
          p.flags |= HGRID_PF_MID_GPE;
          p.vext = sa->vext; /* g = 0: V = vext */
          p.rho = NULL;
        } else {
          p.flags |= HGRID_PF_MID_POT;
          p.pot = sa->pot;
        }
        if ((stage == 1 || stage == 2) && sa->want_post_norm) {
          /* the step's output is this pass's: sum |S|^2 of the stored transform (n2 times the
           * real-space sum, which the finalize multiplies in) */
          p.flags |= HGRID_PF_POST_RED;
          p.partials = part9;
        }
        lg_run_pass(&p);
      } else {
        lg_pass_init(&p, &s, prec, psi, psi, 2);
        p.dir1 = +1;
        lg_run_pass(&p);
        lg_pass_init(&p, &s, prec, psi, psi, 2);
        p.dir1 = -1;
        lg_run_pass(&p);
      }
    }
  }
  /* last: inverse along the contiguous axis (1/N already applied with K), unless deferred */
  long nbpost = nb2;
  if (!defer) {
    lg_pass_init(&p, &s, prec, psi, psi, 2);
    p.dir1 = +1;
    if (sa->want_post_norm) { p.flags |= HGRID_PF_POST_RED; p.partials = part9; }
    if (sa->pot6) {
      /* the inverse transform, then V(t/6) */
      p.flags |= HGRID_PF_MID_POT | HGRID_PF_POT_NOFFT2;
      p.pot = sa->pot6;
      p.c_re = sa->c6_re;
      p.c_im = sa->c6_im;
      p.cons = sa->cons;
      p.mid = 1.0;
    }
    lg_run_pass(&p);
  } else {
    nbpost = (stage == 1 || stage == 2) ? nb2 : nb1;
  }
  if (sa->want_post_norm) {
    hgrid_reduce_job j = {part9, nbpost, HGRID_RED_POST, 1, d->dscal + GS_RED + HGRID_RED_POST};
    fin.job[fin.njobs++] = j;
  }
  if (sa->renorm_target > 0.0) {
    fin.scale = d->dscal + GS_SCALE;
    fin.norm_sum = d->dscal + GS_RED + HGRID_RED_POST;
    fin.target = sa->renorm_target;
    fin.dv = lg_step_factor(g->nx, g->ny, g->nz, g->step) * (defer ? (double)s.n[2] : 1.0);
  } else if (scale) {
    fin.scale = d->dscal + GS_SCALE; /* the pending scale was applied by the first pass */
  }
  if (fin.njobs || fin.scale) lg_hip(hgrid_step_finalize(&fin), "step_finalize");
  d->valid = GV_DEV;
  d->stamp++;
  d->pending = sa->renorm_target > 0.0;
  d->zpend = defer;
  d->kpend = stage == 1 || stage == 2;
}

/* the propagator branches of grid_wf_propagate (reference src/grid_wf.c:390-525) */
static int uses_cutoff(const wf *gwf) {
  return gwf->propagator == WF_2ND_ORDER_CFFT || gwf->propagator == WF_4TH_ORDER_CFFT;
}
static int fourth_order(const wf *gwf) {
  return gwf->propagator == WF_4TH_ORDER_FFT || gwf->propagator == WF_4TH_ORDER_CFFT;
}

static void kinetic_unfused(wf *gwf, REAL complex time) {
  cgrid_fft(gwf->grid);
  if (uses_cutoff(gwf)) grid_wf_propagate_kinetic_cfft(gwf, time);
  else grid_wf_propagate_kinetic_fft(gwf, time);
  cgrid_inverse_fft_norm(gwf->grid);
}

/* unfused sequence for grids with fewer than two non-trivial axes (reference order):
 * K(t/2) V(t) K(t/2), or V(t/6) K(t/2) V(2t/3) K(t/2) V(t/6) for the 4th-order branches */
static void split_step_unfused(wf *gwf, cgrid *potential, REAL complex time, REAL cons) {
  if (fourth_order(gwf)) {
    grid_wf_propagate_potential(gwf, time / 6.0, potential, cons);
    kinetic_unfused(gwf, 0.5 * time);
    grid_wf_propagate_potential(gwf, 2.0 * time / 3.0, potential, cons);
    kinetic_unfused(gwf, 0.5 * time);
    grid_wf_propagate_potential(gwf, time / 6.0, potential, cons);
    return;
  }
  kinetic_unfused(gwf, 0.5 * time);
  grid_wf_propagate_potential(gwf, time, potential, cons);
  kinetic_unfused(gwf, 0.5 * time);
}

/* the propagators the split-operator kernels implement */
static void check_split_propagator(const wf *gwf, const char *who, int allow_fourth) {
  const char p = gwf->propagator;
  if (p == WF_2ND_ORDER_FFT || p == WF_2ND_ORDER_CFFT) return;
  if (allow_fourth && (p == WF_4TH_ORDER_FFT || p == WF_4TH_ORDER_CFFT)) return;
  lg_fatal("Error in %s(). Propagator %d is not implemented by this build (%s).", who, (int)p,
           allow_fourth ? "FFT propagators 1-4 are" : "WF_2ND_ORDER_FFT and WF_2ND_ORDER_CFFT are");
}

static void check_fft_grid(cgrid *g) {
  INT d[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; ++a)
    if (d[a] > 1 && !hgrid_fft_supported(d[a]))
      lg_fatal("FFT length %ld is not supported (powers of two up to 1024, 3*2^k up to 384)", (long)d[a]);
}

static void propagate_split(wf *gwf, cgrid *potential, REAL complex time, int real_space);

void grid_wf_propagate(wf *gwf, cgrid *potential, REAL complex time) {
  cgrid *g = gwf->grid;
  check_split_propagator(gwf, "grid_wf_propagate", 1);
  if (g->omega != 0.0) lg_fatal("omega != 0.0 allowed only with CN.");
  if (gwf->ts_func && fourth_order(gwf)) lg_fatal("4th order propagator not available for FFT absorbing BC.");
  if (gwf->ts_func && creal(time) != 0.0) {
    /* absorbing boundary (reference src/grid_wf.c:396-422): a real-time step of psi and an
     * imaginary-time step (time -i Re t) of a copy, merged with the boundary amplitude; pure
     * imaginary time skips the boundary */
    if (!gwf->cworkspace)
      gwf->cworkspace = cgrid_alloc(g->nx, g->ny, g->nz, g->step, g->value_outside, g->outside_params_ptr, "WF cworkspace 1");
    cgrid *w = gwf->cworkspace;
    if (cgrid_get_precision(w) != cgrid_get_precision(g)) cgrid_set_precision(w, cgrid_get_precision(g));
    cgrid_copy(w, g);
    REAL (*save)(INT, INT, INT, INT, INT, INT, INT, INT, INT) = gwf->ts_func;
    gwf->ts_func = NULL;
    propagate_split(gwf, potential, time, 1);
    gwf->grid = w;
    propagate_split(gwf, potential, -I * creal(time), 1);
    gwf->grid = g;
    gwf->ts_func = save;
    grid_wf_merge(gwf, g, w);
    return;
  }
  propagate_split(gwf, potential, time, 0);
}

/* one split-operator step of gwf->grid by the propagator's sequence; real_space leaves the
 * result untransformed (the absorbing-boundary merge reads it) */
static void propagate_split(wf *gwf, cgrid *potential, REAL complex time, int real_space) {
  cgrid *g = gwf->grid;
  const REAL kx0 = g->kx0, ky0 = g->ky0, kz0 = g->kz0;
  const REAL cons = -(HBAR * HBAR / (2.0 * gwf->mass)) * (kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
  check_fft_grid(g);
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  if (s.m < 2) {
    split_step_unfused(gwf, potential, time, cons);
    return;
  }
  step_args sa;
  memset(&sa, 0, sizeof(sa));
  if (potential) {
    if (potential->nx != g->nx || potential->ny != g->ny || potential->nz != g->nz)
      lg_fatal("grid_wf_propagate: potential and wavefunction grids differ in size");
    /* the potential keeps its own precision (a copy in the state's storage precision is used),
     * so an FP64 state propagated with it later still sees the FP64 values */
    sa.pot = lg_c_in_as(potential, g->dev->prec);
  }
  /* 4th order (reference src/grid_wf.c:433-457): the outer V(t/6) stages ride on the first
   * and last z passes of a fused K(t/2) V(2t/3) K(t/2).  The reference also forms
   * (1/48)(hbar^2/m)|t|^2 |grad V|^2 in a workspace (:449-451) but hands the unmodified
   * potential to the V(2t/3) stage (:452); that term never reaches psi, so it is not formed. */
  const int fourth = fourth_order(gwf);
  const REAL complex tv = fourth ? 2.0 * time / 3.0 : time;
  if (fourth && sa.pot) {
    const REAL complex t6 = time / 6.0;
    sa.pot6 = sa.pot;
    sa.c6_re = cimag(t6) / HBAR;
    sa.c6_im = -creal(t6) / HBAR;
  }
  sa.c_re = cimag(tv) / HBAR;
  sa.c_im = -creal(tv) / HBAR;
  sa.cons = cons;
  const REAL complex tm = -I * (0.5 * time) * HBAR / (gwf->mass * 2.0);
  sa.tm_re = creal(tm);
  sa.tm_im = cimag(tm);
  sa.kcut = uses_cutoff(gwf);
  sa.kcut2 = gwf->kmax * gwf->kmax;
  sa.no_defer = fourth || real_space; /* the last V(t/6) stage and the merge act in real space */
  fused_step(gwf, &sa);
}

/* ---- energies ----------------------------------------------------------------------------- */

/* sum k^2 |FFT(psi)|^2 without modifying psi: forward passes into the workspace, the last one
 * reduces instead of storing. */
static double sum_k2_spectrum(cgrid *g) {
  grid_dev *d = g->dev;
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  check_fft_grid(g);
  if (s.m == 0) return 0.0;
  const double *scale;
  void *psi = lg_c_in(g, &scale);
  void *work = wf_work(g);
  ktab_cache *k2 = lg_k2_tables(g);
  const int aK = 3 - s.m;
  const long nbK = hgrid_pass_blocks(d->prec, s.n[0], s.n[1], s.n[2], aK);
  double *part = lg_partials(HGRID_RED_NCHAN * nbK);
  hgrid_pass p;
  int first = 1;
  for (int axis = 2; axis >= aK; --axis) {
    lg_pass_init(&p, &s, d->prec, first ? psi : work, work, axis);
    if (first && scale) { p.flags |= HGRID_PF_PRESCALE; p.scale = scale; }
    if (axis == aK) {
      p.flags |= HGRID_PF_MID_KRED | HGRID_PF_NOSTORE;
      set_k2(&p, k2, &s, g);
      p.partials = part;
    }
    lg_run_pass(&p);
    first = 0;
  }
  lg_hip(hgrid_reduce_partials(d->dscal + GS_TMP, part + HGRID_RED_KIN * nbK, nbK, 1), "reduce");
  return fetch(d->dscal + GS_TMP);
}

REAL grid_wf_kinetic_energy_fft(wf *gwf) {
  cgrid *g = gwf->grid;
  /* -hbar^2/2m * laplace expectation = hbar^2/2m * sum k^2 |psi_k|^2 * fft_norm2 */
  return HBAR * HBAR / (2.0 * gwf->mass) * sum_k2_spectrum(g) * g->fft_norm2;
}

REAL grid_wf_potential_energy(wf *gwf, rgrid *potential) { return grid_grid_expectation_value(gwf->grid, potential); }

REAL grid_wf_energy_fft(wf *gwf, rgrid *potential) {
  REAL en = grid_wf_kinetic_energy_fft(gwf);
  if (potential) en += grid_wf_potential_energy(gwf, potential);
  return en;
}

/* the boundary rule of src/cgrid_bc.c:21-86 as a kernel code; *dir = the Dirichlet value */
static int bc_code(cgrid *g, REAL complex *dir) {
  *dir = 0.0;
  if (g->value_outside == CGRID_PERIODIC_BOUNDARY) return 2;
  if (g->value_outside == CGRID_NEUMANN_BOUNDARY) return 1;
  if (g->value_outside == CGRID_DIRICHLET_BOUNDARY) {
    *dir = *((REAL complex *)g->outside_params_ptr);
    return 0;
  }
  lg_fatal("user-defined boundary functions are not supported on the GPU");
}

REAL grid_wf_kinetic_energy_cn(wf *gwf) {
  cgrid *g = gwf->grid;
  const double *scale;
  const void *psi = lg_c_in(g, &scale);
  grid_dev *d = g->dev;
  double *scr = lg_partials(hgrid_reduce_scratch((long)d->n));
  REAL complex dir;
  const int bc = bc_code(g, &dir);
  lg_hip(hgrid_fd_laplace_expect(d->dscal + GS_TMP, psi, d->prec, g->nx, g->ny, g->nz, bc, creal(dir), cimag(dir), scale,
                                 scr),
         "fd_laplace_expect");
  /* Re int psi^* (-hbar^2/2m) laplace(psi) (reference src/grid_wf_cn.c:37-49) */
  return -HBAR * HBAR / (2.0 * gwf->mass) * fetch(d->dscal + GS_TMP) / (g->step * g->step) *
         lg_step_factor(g->nx, g->ny, g->nz, g->step);
}

REAL grid_wf_energy_cn(wf *gwf, rgrid *potential) {
  REAL en = grid_wf_kinetic_energy_cn(gwf);
  if (potential) en += grid_wf_potential_energy(gwf, potential);
  return en;
}

REAL grid_wf_energy(wf *gwf, rgrid *potential) {
  if (gwf->propagator == WF_2ND_ORDER_CN || gwf->propagator == WF_4TH_ORDER_CN || !grid_analyze_method)
    return grid_wf_energy_cn(gwf, potential);
  return grid_wf_energy_fft(gwf, potential);
}

REAL grid_wf_kinetic_energy(wf *gwf) { return grid_wf_energy(gwf, NULL); }

/* ---- Gross-Pitaevskii --------------------------------------------------------------------- */

static double gpe_energy_from_parts(wf *gwf, const double *parts, double g_nl) {
  cgrid *g = gwf->grid;
  const double dv = lg_step_factor(g->nx, g->ny, g->nz, g->step);
  const double t = HBAR * HBAR / (2.0 * gwf->mass) * parts[HGRID_RED_KIN] * g->fft_norm2;
  return t + parts[HGRID_RED_VEXT] * dv + 0.5 * g_nl * parts[HGRID_RED_QUART] * dv;
}

static double g_last_g = 0.0;

static void gpe_step(wf *gwf, rgrid *vext, REAL g_nl, REAL complex time, char normalize, int want_energy);

void grid_wf_gpe_propagate(wf *gwf, rgrid *vext, REAL g_nl, REAL complex time, char normalize) {
  gpe_step(gwf, vext, g_nl, time, normalize, 1);
}

/* one fused GPE step; without want_energy the energy reductions are left out, and with
 * g = 0 a deferred 3-D state is not read back at all before the transforms */
static void gpe_stage(wf *gwf, rgrid *vext, REAL g_nl, REAL complex time, char normalize, int want_energy, int stage);

static void gpe_step(wf *gwf, rgrid *vext, REAL g_nl, REAL complex time, char normalize, int want_energy) {
  gpe_stage(gwf, vext, g_nl, time, normalize, want_energy, 0);
}

static void gpe_stage(wf *gwf, rgrid *vext, REAL g_nl, REAL complex time, char normalize, int want_energy, int stage) {
  cgrid *g = gwf->grid;
  check_split_propagator(gwf, "grid_wf_gpe_propagate", 0);
  if (g->omega != 0.0) lg_fatal("omega != 0.0 allowed only with CN.");
  if (vext->nx != g->nx || vext->ny != g->ny || vext->nz != g->nz) lg_fatal("grid_wf_gpe_propagate: size mismatch");
  check_fft_grid(g);
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  if (s.m < 2) lg_fatal("grid_wf_gpe_propagate needs at least two non-trivial axes");
  const REAL kx0 = g->kx0, ky0 = g->ky0, kz0 = g->kz0;
  step_args sa;
  memset(&sa, 0, sizeof(sa));
  sa.gpe = 1;
  sa.vext = lg_r_in(vext);
  sa.g = g_nl;
  sa.c_re = cimag(time) / HBAR;
  sa.c_im = -creal(time) / HBAR;
  sa.cons = -(HBAR * HBAR / (2.0 * gwf->mass)) * (kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
  const REAL complex tm = -I * (0.5 * time) * HBAR / (gwf->mass * 2.0);
  sa.tm_re = creal(tm);
  sa.tm_im = cimag(tm);
  sa.kcut = uses_cutoff(gwf);
  sa.kcut2 = gwf->kmax * gwf->kmax;
  sa.want_energy = want_energy;
  sa.want_post_norm = normalize ? 1 : 0;
  sa.renorm_target = normalize ? gwf->norm : 0.0;
  sa.stage = stage;
  fused_step(gwf, &sa);
  grid_dev *d = g->dev;
  if (stage <= 1) d->have_gpe_energy = want_energy;
  g_last_g = g_nl;
}

REAL grid_wf_gpe_last_energy(wf *gwf) {
  grid_dev *d = gwf->grid->dev;
  if (!d->have_gpe_energy) lg_fatal("grid_wf_gpe_last_energy: no GPE step has run on this wavefunction");
  double parts[HGRID_RED_NCHAN];
  lg_hip(hgrid_memcpy_d2h(parts, d->dscal + GS_RED, sizeof(parts)), "hipMemcpy D2H (energies)");
  return gpe_energy_from_parts(gwf, parts, g_last_g);
}

REAL grid_wf_gpe_energy(wf *gwf, rgrid *vext, REAL g_nl) {
  cgrid *g = gwf->grid;
  const double dv = lg_step_factor(g->nx, g->ny, g->nz, g->step);
  double t = grid_wf_kinetic_energy_fft(gwf);
  double v = grid_grid_expectation_value(g, vext);
  /* (g/2) int |psi|^4 */
  grid_dev *d = g->dev;
  const double *scale;
  const void *psi = lg_c_in(g, &scale);
  if (!d->rho) lg_hip(hgrid_malloc((void **)&d->rho, d->n * sizeof(double)), "hipMalloc (density)");
  lg_hip(hgrid_density(d->rho, psi, d->prec, (long)d->n, scale), "density");
  double *scr = lg_partials(hgrid_reduce_scratch((long)d->n));
  lg_hip(hgrid_sum_real(d->dscal + GS_TMP, d->rho, (long)d->n, 1, scr), "sum_real");
  double q = fetch(d->dscal + GS_TMP);
  return t + v + 0.5 * g_nl * q * dv;
}

static int g_use_graphs = 1;

void grid_hip_graphs(int enable) { g_use_graphs = enable ? 1 : 0; }

static int g_merge_kinetic = 1;

void grid_hip_merged_relaxation(int enable) { g_merge_kinetic = enable ? 1 : 0; }

/* Linear relaxation (g = 0) of a 3-D grid with the plain 2nd-order propagator runs its batches
 * with merged kinetic half-steps: the last K(t/2) of one step and the first of the next are one
 * K(t), so a step costs the four passes y, x (K), y, z (V) instead of seven.  The sequence of
 * states at the batch boundaries, where energies, snapshots and the stopping rule look, is that
 * of whole steps (K K = K(t) in the phases). */
static int merged_batches(const wf *gwf, REAL g_nl) {
  const cgrid *g = gwf->grid;
  lg_shape s;
  lg_squeeze(g->nx, g->ny, g->nz, &s);
  return g_merge_kinetic && g_nl == 0.0 && s.m == 3 && gwf->propagator == WF_2ND_ORDER_FFT;
}

/* Vector extrapolation of the relaxation (used when stride > 0).  Snapshots psi_k are taken
 * every `stride` steps.  With the differences d1 = psi_k - psi_{k-1} and d2 = psi_{k+1} - psi_k,
 * a slowest mode decaying by r per stride shows as d2 = r d1.  When d1 and d2 are parallel
 * (|<d1|d2>| > LG_EXTRAP_COS |d1| |d2|) and r = Re<d1|d2> / |d1|^2 lies in (0.05, 0.995), the
 * mode's remaining geometric tail d2 r / (1 - r) is added at once (vector Aitken), the state is
 * renormalised and the stopping rule restarts (for FP64 grids without its small-rise exit,
 * so a tail that an overshoot left rising is followed to the end).  The fixed point is
 * unchanged; the plain rule still decides convergence. */
#define LG_EXTRAP_COS 0.999

typedef struct {
  void *s1, *d1, *d2;  /* device arrays of the grid's precision: last snapshot, last difference, scratch */
  int have;            /* snapshots since the last extrapolation (0, 1, 2+) */
  long count;          /* extrapolations done */
} extrap_state;

static double complex extrap_dot(cgrid *g, const void *a, const void *b) {
  grid_dev *d = g->dev;
  double *scr = lg_partials(hgrid_reduce_scratch((long)d->n));
  lg_hip(hgrid_sum_conj_product(d->dscal + GS_TMP, a, b, d->prec, (long)d->n, scr), "sum_conj_product");
  double v[2];
  lg_hip(hgrid_memcpy_d2h(v, d->dscal + GS_TMP, sizeof(v)), "hipMemcpy D2H (scalar)");
  return v[0] + I * v[1];
}

/* returns 1 if it changed the state; *r_out receives the ratio per stride it used */
static int extrap_snapshot(wf *gwf, extrap_state *x, double *r_out) {
  cgrid *g = gwf->grid;
  grid_dev *d = g->dev;
  const size_t bytes = d->n * lg_prec_size(d->prec);
  const long n = (long)d->n;
  if (!x->s1) {
    lg_hip(hgrid_malloc(&x->s1, bytes), "hipMalloc (relax snapshot)");
    lg_hip(hgrid_malloc(&x->d1, bytes), "hipMalloc (relax difference)");
    lg_hip(hgrid_malloc(&x->d2, bytes), "hipMalloc (relax difference)");
  }
  void *psi = lg_c_in_flushed(g); /* real space, no pending scale */
  if (x->have == 0) {
    lg_hip(hgrid_memcpy_d2d(x->s1, psi, bytes), "hipMemcpy D2D (snapshot)");
    x->have = 1;
    return 0;
  }
  lg_hip(hgrid_memcpy_d2d(x->d2, psi, bytes), "hipMemcpy D2D (snapshot)");
  lg_hip(hgrid_caxpy(x->d2, x->s1, d->prec, n, -1.0, 0.0), "caxpy");
  lg_hip(hgrid_memcpy_d2d(x->s1, psi, bytes), "hipMemcpy D2D (snapshot)");
  void *t = x->d1;
  if (x->have == 1) {
    x->d1 = x->d2;
    x->d2 = t;
    x->have = 2;
    return 0;
  }
  const double n11 = creal(extrap_dot(g, x->d1, x->d1));
  const double n22 = creal(extrap_dot(g, x->d2, x->d2));
  const double complex x12 = extrap_dot(g, x->d1, x->d2);
  if (!isfinite(n11) || !isfinite(n22) || !isfinite(creal(x12)) || !isfinite(cimag(x12)))
    lg_fatal("imaginary-time relaxation diverged: non-finite snapshot differences at an extrapolation check");
  const double r = creal(x12) / n11;
  if (n11 > 0.0 && n22 > 0.0 && cabs(x12) > LG_EXTRAP_COS * sqrt(n11 * n22) && r > 0.05 && r < 0.995) {
    *r_out = r;
    lg_hip(hgrid_caxpy(psi, x->d2, d->prec, n, r / (1.0 - r), 0.0), "caxpy");
    lg_c_written(g, 1);
    grid_wf_normalize(gwf);
    x->have = 0;
    x->count++;
    return 1;
  }
  x->d1 = x->d2;
  x->d2 = t;
  return 0;
}

/* The device addresses a captured relaxation batch reads or writes.  Tables and the shared
 * reduction scratch may be reallocated by calls made between batches (an extrapolation's
 * reductions, a precision change), so a graph is replayed only while every one of them is
 * unchanged, and is captured again otherwise. */
typedef struct {
  const void *p[17];
  unsigned long partials_gen;
} graph_key;

static void graph_key_of(wf *gwf, rgrid *vext, const double *hist, graph_key *k) {
  grid_dev *d = gwf->grid->dev;
  memset(k, 0, sizeof(*k));
  int i = 0;
  k->p[i++] = d->d;
  k->p[i++] = d->dscal;
  k->p[i++] = d->rho;
  k->p[i++] = vext->dev->d;
  k->p[i++] = hist;
  for (int a = 0; a < 3; ++a) {
    k->p[i++] = d->kin.t[a];
    k->p[i++] = d->kin_energy.k2[a];
  }
  k->p[i++] = d->kin.kc[0];
  k->p[i++] = d->kin.kc[1];
  k->p[i++] = d->kin.kc[2];
  for (int a = 0; a < 3; ++a) k->p[i++] = d->kin2.t[a];
  k->partials_gen = lg_partials_generation();
}

static INT relax_impl(wf *gwf, rgrid *vext, REAL g_nl, REAL dt, REAL tol, INT max_steps, REAL *energy, INT stride,
                      long *n_extrap) {
  grid_dev *d = gwf->grid->dev;
  const INT batch = 16;
#if LG_RELAX_MIN_STRIDE % 16 != 0
#error "the relaxation batches assume stopping-rule strides that are multiples of 16"
#endif
  if (stride < 0 || stride % batch) lg_fatal("grid_wf_gpe_relax_extrapolated: stride %ld is not a multiple of 16", (long)stride);
  double *hist = NULL;
  lg_ensure_device();
  lg_hip(hgrid_malloc((void **)&hist, sizeof(double) * HGRID_RED_NCHAN), "hipMalloc (energy history)");
  double host[HGRID_RED_NCHAN];
  double e = NAN;
  const int merge = merged_batches(gwf, g_nl);
  INT n = 0, n0 = 0; /* n0: step at which the stopping rule (re)started */
  int converged = 0;
  lg_relax_state rs;
  lg_relax_init(&rs);
  extrap_state xs = {NULL, NULL, NULL, 0, 0};
  /* After the first batch every step is the same sequence of launches on the same buffers
   * (tables cached, scale pending, 3-D state deferred), so one batch of 16 steps is captured
   * as a HIP graph and replayed: small grids are launch-bound otherwise.  A snapshot or an
   * extrapolation brings the state back to real space with no pending scale, so the batch
   * after one runs eagerly and restores the captured form. */
  void *graph = NULL;
  graph_key key, now;
  int eager = 1;
  while (n < max_steps && !converged) {
    INT nb = (max_steps - n) < batch ? (max_steps - n) : batch;
    if (graph && nb == batch && !eager) {
      graph_key_of(gwf, vext, hist, &now);
      if (memcmp(&now, &key, sizeof(key))) {
        /* a buffer the graph holds has moved: drop it and run this batch eagerly */
        lg_hip(hgrid_graph_free(graph), "hipGraphExecDestroy");
        graph = NULL;
        eager = 1;
      }
    }
    if (graph && nb == batch && !eager) {
      lg_hip(hgrid_graph_launch(graph), "hipGraphLaunch (relaxation)");
    } else {
      const int capture = g_use_graphs && !graph && !eager && nb == batch;
      if (capture) lg_hip(hgrid_graph_begin(), "hipStreamBeginCapture (relaxation)");
      for (INT b = 0; b < nb; ++b) {
        /* the rule's checkpoints fall on multiples of its strides, all multiples of 16 from
         * its start, so only the first step of a batch computes its input state's energy */
        gpe_stage(gwf, vext, g_nl, -I * dt, 1, b == 0, merge ? (b == 0 ? 1 : 2) : 0);
        if (b == 0) lg_hip(hgrid_copy_f64(hist, d->dscal + GS_RED, HGRID_RED_NCHAN), "copy (energy history)");
      }
      if (merge) gpe_stage(gwf, vext, g_nl, -I * dt, 1, 0, 3);
      if (capture) {
        lg_hip(hgrid_graph_end(&graph), "hipStreamEndCapture (relaxation)");
        graph_key_of(gwf, vext, hist, &key);
        lg_hip(hgrid_graph_launch(graph), "hipGraphLaunch (relaxation)"); /* capture does not execute */
      }
      eager = 0;
    }
    lg_hip(hgrid_memcpy_d2h(host, hist, sizeof(double) * HGRID_RED_NCHAN), "hipMemcpy D2H (energy history)");
    e = gpe_energy_from_parts(gwf, host, g_nl); /* the energy of the state entering step n */
    const int verdict = lg_relax_feed(&rs, (long)(n - n0), e, tol);
    if (verdict < 0)
      lg_fatal("imaginary-time relaxation diverged: the energy of the state entering step %ld is %g", (long)n, e);
    converged = verdict > 0;
    n += nb;
    if (!converged && stride && nb == batch && n % stride == 0) {
      double r = 0.0;
      if (extrap_snapshot(gwf, &xs, &r)) {
        lg_relax_init(&rs);
        /* an overshoot leaves the energy rising towards the limit: in FP64 that is no rounding */
        rs.rise_exit = d->prec != HGRID_C128;
        /* what is left of the slowest mode decays by r per stride at most */
        rs.lnq_floor = log(r) / (double)stride;
        n0 = n;
      }
      eager = 1;
    }
  }
  if (graph) lg_hip(hgrid_graph_free(graph), "hipGraphExecDestroy");
  lg_hip(hgrid_free(hist), "hipFree");
  if (xs.s1) {
    lg_hip(hgrid_free(xs.s1), "hipFree");
    lg_hip(hgrid_free(xs.d1), "hipFree");
    lg_hip(hgrid_free(xs.d2), "hipFree");
  }
  if (n_extrap) *n_extrap = xs.count;
  /* the energy of the state returned (the stopping rule decided on a checkpoint up to 15
   * steps earlier) */
  if (energy) {
    *energy = grid_wf_gpe_energy(gwf, vext, g_nl);
    if (!isfinite(*energy)) lg_fatal("imaginary-time relaxation diverged: final energy %g", *energy);
  }
  return n;
}

INT grid_wf_gpe_relax(wf *gwf, rgrid *vext, REAL g_nl, REAL dt, REAL tol, INT max_steps, REAL *energy) {
  return relax_impl(gwf, vext, g_nl, dt, tol, max_steps, energy, 0, NULL);
}

INT grid_wf_gpe_relax_extrapolated(wf *gwf, rgrid *vext, REAL g_nl, REAL dt, REAL tol, INT max_steps, REAL *energy,
                                   INT stride, INT *n_extrap) {
  long ne = 0;
  const INT steps = relax_impl(gwf, vext, g_nl, dt, tol, max_steps, energy, stride, &ne);
  if (n_extrap) *n_extrap = ne;
  return steps;
}
