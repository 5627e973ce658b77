/*
 * Learned warm start for imaginary-time relaxation: a 3-D U-Net that maps a potential sampled
 * on an S^3 grid (S = 64 for the 256^3 problems) to a ground-state amplitude, upsampled to the
 * wavefunction's grid as the start of grid_wf_gpe_relax.  The reference has no learned
 * component; the network's forward and backward kernels are hip/nn.hip, this file holds the
 * parameters, the layer schedule, Adam and the training loop.  With no network the start is
 * the plain Gaussian (grid_wf_gaussian_start), and the relaxation itself always decides the
 * final energy.
 *
 * Architecture (channels; ReLU after every 3x3x3 periodic convolution; levels S .. S/8):
 *   e1 8->16 | pool | e2a 16->32, e2b 32->32 | pool | e3a 32->64, e3b 64->64 | pool |
 *   ba 64->128, bb 128->128 | up, concat e3 | d3a 192->64, d3b 64->64 | up, concat e2 |
 *   d2a 96->32, d2b 32->32 | up, concat e1 | d1 48->16 | 1x1 head 16->1.
 * Skip tensors are written straight into the upper part of the concat buffers (c1, c2, c3).
 * Input channel 0 is vscale * V, channels 1..7 are zero.  Parameters are one FP32 array in
 * layer order (W [Cout][27][Cin] then b [Cout]), then the head (16 weights, 1 bias).
 */
#define _GNU_SOURCE
#include <math.h>

#include "internal.h"

typedef struct {
  const char *name;
  int cin, cout;
} layer_def;

enum { E1, E2A, E2B, E3A, E3B, BA, BB, D3A, D3B, D2A, D2B, D1, NLAYERS };
static const layer_def kL[NLAYERS] = {{"e1", 8, 16},    {"e2a", 16, 32},  {"e2b", 32, 32},   {"e3a", 32, 64},
                                      {"e3b", 64, 64},  {"ba", 64, 128},  {"bb", 128, 128},  {"d3a", 192, 64},
                                      {"d3b", 64, 64},  {"d2a", 96, 32},  {"d2b", 32, 32},   {"d1", 48, 16}};

struct grid_surrogate {
  int lgS, batch;
  long nv; /* S^3 */
  float vscale;
  long np, off_w[NLAYERS], off_b[NLAYERS], off_head;
  float *w, *g, *m1, *m2; /* device FP32 parameters, gradients, Adam moments */
  void *wb[NLAYERS], *wt[NLAYERS];
  int adam_t;
  /* activations (bf16), sized for batch samples */
  void *x0, *c1, *p1, *e2a, *c2, *p2, *e3a, *c3, *p3, *ba, *bb, *d3a, *d3, *d2a, *d2, *d1;
  /* gradients (bf16) */
  void *gc1, *gc2, *gc3, *gt1, *gt2;
  float *vin, *tgt, *out, *partial, *bpartial, *scal;
  double *clip_scratch;
  long partial_n;
  /* device-resident training set */
  float *set_v, *set_t;
  long set_n;
};

static void *dalloc(size_t bytes, const char *what) {
  void *p;
  lg_hip(hgrid_malloc(&p, bytes), what);
  return p;
}

#define BF(n) ((size_t)(n) * 2)

static long lvl(const grid_surrogate *s, long nvb, int l) {
  (void)s;
  return nvb >> (3 * l);
}

/* host copy of W rows for the BF16 caches */
static void refresh_weights(grid_surrogate *s) {
  for (int l = 0; l < NLAYERS; ++l)
    lg_hip(hgrid_nn_wprep(s->w + s->off_w[l], kL[l].cout, kL[l].cin, s->wb[l], s->wt[l]), "nn_wprep");
}

grid_surrogate *grid_surrogate_alloc(INT side, INT batch, REAL vscale, INT seed) {
  int lgS = 0;
  while ((1L << lgS) < side) ++lgS;
  if ((1L << lgS) != side || lgS < 4) lg_fatal("grid_surrogate_alloc: side must be a power of two >= 16");
  if (batch < 1) lg_fatal("grid_surrogate_alloc: batch must be >= 1");
  lg_ensure_device();
  grid_surrogate *s = (grid_surrogate *)calloc(1, sizeof(grid_surrogate));
  if (!s) lg_fatal("out of memory (surrogate)");
  s->lgS = lgS;
  s->batch = (int)batch;
  s->nv = 1L << (3 * lgS);
  s->vscale = (float)vscale;
  long o = 0;
  for (int l = 0; l < NLAYERS; ++l) {
    s->off_w[l] = o;
    o += 27L * kL[l].cin * kL[l].cout;
    s->off_b[l] = o;
    o += kL[l].cout;
  }
  s->off_head = o;
  o += 17;
  s->np = o;
  s->w = (float *)dalloc(sizeof(float) * o, "hipMalloc (surrogate params)");
  s->g = (float *)dalloc(sizeof(float) * o, "hipMalloc (surrogate grads)");
  s->m1 = (float *)dalloc(sizeof(float) * o, "hipMalloc (adam)");
  s->m2 = (float *)dalloc(sizeof(float) * o, "hipMalloc (adam)");
  lg_hip(hgrid_memset(s->m1, 0, sizeof(float) * o), "memset");
  lg_hip(hgrid_memset(s->m2, 0, sizeof(float) * o), "memset");
  for (int l = 0; l < NLAYERS; ++l) {
    s->wb[l] = dalloc(BF(27L * kL[l].cin * kL[l].cout), "hipMalloc (bf16 weights)");
    s->wt[l] = dalloc(BF(27L * kL[l].cin * kL[l].cout), "hipMalloc (bf16 weights)");
  }
  /* He initialisation from a seeded SplitMix64 / Box-Muller stream */
  double *h = (double *)malloc(sizeof(double) * o);
  unsigned long long st = 0x9e3779b97f4a7c15ULL * (unsigned long long)(seed + 1);
  for (int l = 0; l < NLAYERS; ++l) {
    const double sd = sqrt(2.0 / (27.0 * kL[l].cin));
    for (long i = s->off_w[l]; i < s->off_b[l]; ++i) h[i] = sd * lg_normal(&st);
    for (long i = s->off_b[l]; i < s->off_b[l] + kL[l].cout; ++i) h[i] = 0.0;
  }
  for (int i = 0; i < 16; ++i) h[s->off_head + i] = 0.25 * lg_normal(&st);
  h[s->off_head + 16] = 0.0;
  grid_surrogate_set_params(s, h);
  free(h);

  const long nvb = s->nv * batch;
  s->x0 = dalloc(BF(nvb * 8), "hipMalloc (surrogate activations)");
  s->c1 = dalloc(BF(nvb * 48), "hipMalloc (surrogate activations)");
  s->d1 = dalloc(BF(nvb * 16), "hipMalloc (surrogate activations)");
  const long n1 = lvl(s, nvb, 1), n2 = lvl(s, nvb, 2), n3 = lvl(s, nvb, 3);
  s->p1 = dalloc(BF(n1 * 16), "hipMalloc (surrogate activations)");
  s->e2a = dalloc(BF(n1 * 32), "hipMalloc (surrogate activations)");
  s->c2 = dalloc(BF(n1 * 96), "hipMalloc (surrogate activations)");
  s->d2a = dalloc(BF(n1 * 32), "hipMalloc (surrogate activations)");
  s->d2 = dalloc(BF(n1 * 32), "hipMalloc (surrogate activations)");
  s->p2 = dalloc(BF(n2 * 32), "hipMalloc (surrogate activations)");
  s->e3a = dalloc(BF(n2 * 64), "hipMalloc (surrogate activations)");
  s->c3 = dalloc(BF(n2 * 192), "hipMalloc (surrogate activations)");
  s->d3a = dalloc(BF(n2 * 64), "hipMalloc (surrogate activations)");
  s->d3 = dalloc(BF(n2 * 64), "hipMalloc (surrogate activations)");
  s->p3 = dalloc(BF(n3 * 64), "hipMalloc (surrogate activations)");
  s->ba = dalloc(BF(n3 * 128), "hipMalloc (surrogate activations)");
  s->bb = dalloc(BF(n3 * 128), "hipMalloc (surrogate activations)");
  s->gc1 = dalloc(BF(nvb * 48), "hipMalloc (surrogate gradients)");
  s->gc2 = dalloc(BF(n1 * 96), "hipMalloc (surrogate gradients)");
  s->gc3 = dalloc(BF(n2 * 192), "hipMalloc (surrogate gradients)");
  s->gt1 = dalloc(BF(nvb * 16), "hipMalloc (surrogate gradients)");
  s->gt2 = dalloc(BF(nvb * 16), "hipMalloc (surrogate gradients)");
  s->vin = (float *)dalloc(sizeof(float) * nvb, "hipMalloc (surrogate input)");
  s->tgt = (float *)dalloc(sizeof(float) * nvb, "hipMalloc (surrogate target)");
  s->out = (float *)dalloc(sizeof(float) * nvb, "hipMalloc (surrogate output)");
  /* partial-sum scratch: the largest of the weight-gradient, bias and head partials */
  long pn = 0;
  for (int l = 0; l < NLAYERS; ++l) {
    const int lev = l == E1 || l == D1 ? 0 : (l == E2A || l == E2B || l == D2A || l == D2B) ? 1
                                          : (l == E3A || l == E3B || l == D3A || l == D3B) ? 2 : 3;
    const long nvl = lvl(s, nvb, lev);
    const long a = hgrid_nn_wgrad_partials(nvl) * 27L * kL[l].cin * kL[l].cout;
    const long bsum = hgrid_nn_colsum_chunks(nvl) * kL[l].cout;
    if (a > pn) pn = a;
    if (bsum > pn) pn = bsum;
  }
  if (hgrid_nn_head_blocks(nvb) * 18 > pn) pn = hgrid_nn_head_blocks(nvb) * 18;
  s->partial_n = pn;
  s->partial = (float *)dalloc(sizeof(float) * pn, "hipMalloc (surrogate partials)");
  s->scal = (float *)dalloc(sizeof(float) * 4, "hipMalloc (surrogate scalars)");
  s->clip_scratch = (double *)dalloc(sizeof(double) * 512, "hipMalloc (clip scratch)");
  s->bpartial = (float *)dalloc(sizeof(float) * (size_t)(hgrid_nn_wgrad_chunks(nvb) * 192), "hipMalloc (bias partials)");
  return s;
}

void grid_surrogate_free(grid_surrogate *s) {
  if (!s) return;
  void *bufs[] = {s->w,   s->g,   s->m1,  s->m2,  s->x0,  s->c1,  s->p1,  s->e2a, s->c2,  s->p2,  s->e3a,
                  s->c3,  s->p3,  s->ba,  s->bb,  s->d3a, s->d3,  s->d2a, s->d2,  s->d1,  s->gc1, s->gc2,
                  s->gc3, s->gt1, s->gt2, s->vin, s->tgt, s->out, s->partial, s->bpartial, s->scal, s->set_v, s->set_t, s->clip_scratch};
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i)
    if (bufs[i]) lg_hip(hgrid_free(bufs[i]), "hipFree");
  for (int l = 0; l < NLAYERS; ++l) {
    lg_hip(hgrid_free(s->wb[l]), "hipFree");
    lg_hip(hgrid_free(s->wt[l]), "hipFree");
  }
  free(s);
}

INT grid_surrogate_nparams(grid_surrogate *s) { return s->np; }
INT grid_surrogate_side(grid_surrogate *s) { return 1L << s->lgS; }

void grid_surrogate_set_params(grid_surrogate *s, const REAL *p) {
  float *h = (float *)malloc(sizeof(float) * s->np);
  for (long i = 0; i < s->np; ++i) h[i] = (float)p[i];
  lg_hip(hgrid_memcpy_h2d(s->w, h, sizeof(float) * s->np), "hipMemcpy (params)");
  free(h);
  refresh_weights(s);
  s->adam_t = 0;
  lg_hip(hgrid_memset(s->m1, 0, sizeof(float) * s->np), "memset");
  lg_hip(hgrid_memset(s->m2, 0, sizeof(float) * s->np), "memset");
}

void grid_surrogate_get_params(grid_surrogate *s, REAL *p) {
  float *h = (float *)malloc(sizeof(float) * s->np);
  lg_hip(hgrid_memcpy_d2h(h, s->w, sizeof(float) * s->np), "hipMemcpy (params)");
  for (long i = 0; i < s->np; ++i) p[i] = h[i];
  free(h);
}

/* ---- forward / backward ------------------------------------------------------------------- */

static void conv(grid_surrogate *s, int l, const void *x, int xs, int xo, void *y, int ys, int yo, int lev, long nvb) {
  lg_hip(hgrid_nn_conv3(x, xs, xo, kL[l].cin, s->wb[l], s->w + s->off_b[l], y, ys, yo, kL[l].cout, NULL, 0, 0,
                        s->lgS - lev, lvl(s, nvb, lev), 1),
         "nn_conv3");
}

static void forward(grid_surrogate *s, long nvb) {
  const int lg = s->lgS;
  lg_hip(hgrid_nn_input(s->vin, s->vscale, s->x0, nvb), "nn_input");
  conv(s, E1, s->x0, 8, 0, s->c1, 48, 32, 0, nvb);
  lg_hip(hgrid_nn_pool2(s->c1, 48, 32, 16, s->p1, lg - 1, lvl(s, nvb, 1)), "nn_pool2");
  conv(s, E2A, s->p1, 16, 0, s->e2a, 32, 0, 1, nvb);
  conv(s, E2B, s->e2a, 32, 0, s->c2, 96, 64, 1, nvb);
  lg_hip(hgrid_nn_pool2(s->c2, 96, 64, 32, s->p2, lg - 2, lvl(s, nvb, 2)), "nn_pool2");
  conv(s, E3A, s->p2, 32, 0, s->e3a, 64, 0, 2, nvb);
  conv(s, E3B, s->e3a, 64, 0, s->c3, 192, 128, 2, nvb);
  lg_hip(hgrid_nn_pool2(s->c3, 192, 128, 64, s->p3, lg - 3, lvl(s, nvb, 3)), "nn_pool2");
  conv(s, BA, s->p3, 64, 0, s->ba, 128, 0, 3, nvb);
  conv(s, BB, s->ba, 128, 0, s->bb, 128, 0, 3, nvb);
  lg_hip(hgrid_nn_up2(s->bb, 128, 0, 128, s->c3, 192, 0, lg - 2, lvl(s, nvb, 2)), "nn_up2");
  conv(s, D3A, s->c3, 192, 0, s->d3a, 64, 0, 2, nvb);
  conv(s, D3B, s->d3a, 64, 0, s->d3, 64, 0, 2, nvb);
  lg_hip(hgrid_nn_up2(s->d3, 64, 0, 64, s->c2, 96, 0, lg - 1, lvl(s, nvb, 1)), "nn_up2");
  conv(s, D2A, s->c2, 96, 0, s->d2a, 32, 0, 1, nvb);
  conv(s, D2B, s->d2a, 32, 0, s->d2, 32, 0, 1, nvb);
  lg_hip(hgrid_nn_up2(s->d2, 32, 0, 32, s->c1, 48, 0, lg, nvb), "nn_up2");
  conv(s, D1, s->c1, 48, 0, s->d1, 16, 0, 0, nvb);
  lg_hip(hgrid_nn_head_fwd(s->d1, s->w + s->off_head, s->out, nvb), "nn_head_fwd");
}

/* weight and bias gradients of layer l from its masked output gradient dy (dense) and input x */
static void wgrad(grid_surrogate *s, int l, const void *dy, const void *x, int xs, int xo, int lev, long nvb) {
  const long nvl = lvl(s, nvb, lev);
  const int ci = kL[l].cin, co = kL[l].cout;
  lg_hip(hgrid_nn_conv3_wgrad(dy, co, 0, co, x, xs, xo, ci, s->lgS - lev, nvl, s->partial, s->bpartial),
         "nn_conv3_wgrad");
  lg_hip(hgrid_nn_reduce(s->partial, hgrid_nn_wgrad_partials(nvl), 27L * ci * co, s->g + s->off_w[l], 0), "nn_reduce");
  lg_hip(hgrid_nn_reduce(s->bpartial, hgrid_nn_wgrad_chunks(nvl), co, s->g + s->off_b[l], 0), "nn_reduce (bias)");
}

/* data gradient of layer l: dx = conv^T(dy), optionally times (act > 0) */
static void dgrad(grid_surrogate *s, int l, const void *dy, void *dx, const void *act, int as, int ao, int lev, long nvb) {
  lg_hip(hgrid_nn_conv3(dy, kL[l].cout, 0, kL[l].cout, s->wt[l], NULL, dx, kL[l].cin, 0, kL[l].cin, act, as, ao,
                        s->lgS - lev, lvl(s, nvb, lev), 0),
         "nn_conv3 (data gradient)");
}

/* loss and all gradients for the batch in vin/tgt (nb samples) */
static double backward(grid_surrogate *s, long nb) {
  const int lg = s->lgS;
  const long nvb = s->nv * nb;
  const long hb = hgrid_nn_head_blocks(nvb);
  lg_hip(hgrid_nn_head_bwd(s->d1, s->w + s->off_head, s->out, s->tgt, (float)(2.0 / (double)nvb), s->gt1, nvb,
                           s->partial),
         "nn_head_bwd");
  /* head partials [block][18]: columns 0..16 are the head gradient, column 17 the squared error */
  lg_hip(hgrid_nn_reduce_strided(s->partial, hb, 18, 17, s->g + s->off_head), "nn_reduce (head)");
  lg_hip(hgrid_nn_reduce_strided(s->partial + 17, hb, 18, 1, s->scal), "nn_reduce (loss)");
  /* d1 */
  wgrad(s, D1, s->gt1, s->c1, 48, 0, 0, nvb);
  dgrad(s, D1, s->gt1, s->gc1, NULL, 0, 0, 0, nvb);
  lg_hip(hgrid_nn_up2_bwd(s->gc1, 48, 0, 32, s->d2, 32, 0, s->gt2, lg - 1, lvl(s, nvb, 1)), "nn_up2_bwd");
  /* d2b, d2a */
  wgrad(s, D2B, s->gt2, s->d2a, 32, 0, 1, nvb);
  dgrad(s, D2B, s->gt2, s->gt1, s->d2a, 32, 0, 1, nvb);
  wgrad(s, D2A, s->gt1, s->c2, 96, 0, 1, nvb);
  dgrad(s, D2A, s->gt1, s->gc2, NULL, 0, 0, 1, nvb);
  lg_hip(hgrid_nn_up2_bwd(s->gc2, 96, 0, 64, s->d3, 64, 0, s->gt2, lg - 2, lvl(s, nvb, 2)), "nn_up2_bwd");
  /* d3b, d3a */
  wgrad(s, D3B, s->gt2, s->d3a, 64, 0, 2, nvb);
  dgrad(s, D3B, s->gt2, s->gt1, s->d3a, 64, 0, 2, nvb);
  wgrad(s, D3A, s->gt1, s->c3, 192, 0, 2, nvb);
  dgrad(s, D3A, s->gt1, s->gc3, NULL, 0, 0, 2, nvb);
  lg_hip(hgrid_nn_up2_bwd(s->gc3, 192, 0, 128, s->bb, 128, 0, s->gt2, lg - 3, lvl(s, nvb, 3)), "nn_up2_bwd");
  /* bottleneck */
  wgrad(s, BB, s->gt2, s->ba, 128, 0, 3, nvb);
  dgrad(s, BB, s->gt2, s->gt1, s->ba, 128, 0, 3, nvb);
  wgrad(s, BA, s->gt1, s->p3, 64, 0, 3, nvb);
  dgrad(s, BA, s->gt1, s->gt2, NULL, 0, 0, 3, nvb);
  /* e3 = c3[128:192]: pooled gradient + skip gradient, masked */
  lg_hip(hgrid_nn_pool2_bwd(s->gt2, s->gc3, 192, 128, 64, s->c3, 192, 128, s->gt1, lg - 2, lvl(s, nvb, 2)), "nn_pool2_bwd");
  wgrad(s, E3B, s->gt1, s->e3a, 64, 0, 2, nvb);
  dgrad(s, E3B, s->gt1, s->gt2, s->e3a, 64, 0, 2, nvb);
  wgrad(s, E3A, s->gt2, s->p2, 32, 0, 2, nvb);
  dgrad(s, E3A, s->gt2, s->gt1, NULL, 0, 0, 2, nvb);
  lg_hip(hgrid_nn_pool2_bwd(s->gt1, s->gc2, 96, 64, 32, s->c2, 96, 64, s->gt2, lg - 1, lvl(s, nvb, 1)), "nn_pool2_bwd");
  wgrad(s, E2B, s->gt2, s->e2a, 32, 0, 1, nvb);
  dgrad(s, E2B, s->gt2, s->gt1, s->e2a, 32, 0, 1, nvb);
  wgrad(s, E2A, s->gt1, s->p1, 16, 0, 1, nvb);
  dgrad(s, E2A, s->gt1, s->gt2, NULL, 0, 0, 1, nvb);
  lg_hip(hgrid_nn_pool2_bwd(s->gt2, s->gc1, 48, 32, 16, s->c1, 48, 32, s->gt1, lg, nvb), "nn_pool2_bwd");
  wgrad(s, E1, s->gt1, s->x0, 8, 0, 0, nvb);
  float lsum;
  lg_hip(hgrid_memcpy_d2h(&lsum, s->scal, sizeof(float)), "hipMemcpy (loss)");
  return (double)lsum / (double)nvb;
}

/* Adam with the gradient clipped to global norm 1: a single bad minibatch (a loss spike) can
 * otherwise push every ReLU of a layer negative at once, after which the network stays dead */
static void adam_step(grid_surrogate *s, double lr) {
  ++s->adam_t;
  lg_hip(hgrid_nn_clip_factor(s->g, s->np, 1.0, s->clip_scratch, s->scal + 2), "nn_clip_factor");
  lg_hip(hgrid_nn_adam(s->w, s->g, s->m1, s->m2, s->np, (float)lr, 0.9f, 0.999f, 1e-8f, s->adam_t, s->scal + 2),
         "nn_adam");
  refresh_weights(s);
}

static void upload_batch(grid_surrogate *s, long nb, const REAL *pot, const REAL *target) {
  const long n = s->nv * nb;
  if (n <= 0) return;
  float *h = (float *)malloc(sizeof(float) * (size_t)n);
  if (!h) lg_fatal("out of host memory (surrogate batch)");
  for (long i = 0; i < n; ++i) h[i] = (float)pot[i];
  lg_hip(hgrid_memcpy_h2d(s->vin, h, sizeof(float) * n), "hipMemcpy (surrogate input)");
  if (target) {
    for (long i = 0; i < n; ++i) h[i] = (float)target[i];
    lg_hip(hgrid_memcpy_h2d(s->tgt, h, sizeof(float) * n), "hipMemcpy (surrogate target)");
  }
  free(h);
}

REAL grid_surrogate_gradient(grid_surrogate *s, INT nb, const REAL *pot, const REAL *target, REAL *grad) {
  if (nb < 1 || nb > s->batch) lg_fatal("grid_surrogate_gradient: 1 <= nb <= batch");
  upload_batch(s, nb, pot, target);
  forward(s, s->nv * nb);
  const double loss = backward(s, nb);
  if (grad) {
    float *h = (float *)malloc(sizeof(float) * s->np);
    lg_hip(hgrid_memcpy_d2h(h, s->g, sizeof(float) * s->np), "hipMemcpy (grads)");
    for (long i = 0; i < s->np; ++i) grad[i] = h[i];
    free(h);
  }
  return loss;
}

void grid_surrogate_predict(grid_surrogate *s, INT n, const REAL *pot, REAL *amp) {
  float *h = (float *)malloc(sizeof(float) * s->nv * s->batch);
  for (INT i0 = 0; i0 < n; i0 += s->batch) {
    const long nb = (n - i0) < s->batch ? (n - i0) : s->batch;
    upload_batch(s, nb, pot + i0 * s->nv, NULL);
    forward(s, s->nv * nb);
    lg_hip(hgrid_memcpy_d2h(h, s->out, sizeof(float) * s->nv * nb), "hipMemcpy (surrogate output)");
    for (long i = 0; i < s->nv * nb; ++i) amp[i0 * s->nv + i] = h[i];
  }
  free(h);
}

/* ---- training ------------------------------------------------------------------------------ */

void grid_surrogate_set_training_data(grid_surrogate *s, INT nsamples, const REAL *pot, const REAL *target) {
  if (s->set_v) lg_hip(hgrid_free(s->set_v), "hipFree");
  if (s->set_t) lg_hip(hgrid_free(s->set_t), "hipFree");
  s->set_n = nsamples;
  const size_t bytes = sizeof(float) * (size_t)(s->nv * nsamples);
  s->set_v = (float *)dalloc(bytes, "hipMalloc (training potentials)");
  s->set_t = (float *)dalloc(bytes, "hipMalloc (training targets)");
  float *h = (float *)malloc(sizeof(float) * s->nv);
  for (INT k = 0; k < nsamples; ++k) {
    for (long i = 0; i < s->nv; ++i) h[i] = (float)pot[k * s->nv + i];
    lg_hip(hgrid_memcpy_h2d(s->set_v + k * s->nv, h, sizeof(float) * s->nv), "hipMemcpy (training set)");
    for (long i = 0; i < s->nv; ++i) h[i] = (float)target[k * s->nv + i];
    lg_hip(hgrid_memcpy_h2d(s->set_t + k * s->nv, h, sizeof(float) * s->nv), "hipMemcpy (training set)");
  }
  free(h);
}

/* Epochs of Adam over the resident training set in shuffled minibatches (Fisher-Yates from a
 * seeded SplitMix64 stream; with augment, each sample drawn under a random symmetry of the
 * periodic cube: axis permutation, reflections, translation); the learning rate rises linearly over the first 5% of the steps
 * (at least 20) and decays by a cosine from lr to lr / 20 over the run.  loss_history (may be NULL) receives each epoch's mean training loss; returns the last. */
REAL grid_surrogate_train(grid_surrogate *s, INT epochs, REAL lr, INT seed, REAL *loss_history) {
  return grid_surrogate_train_ex(s, epochs, lr, seed, 0, loss_history);
}

REAL grid_surrogate_train_ex(grid_surrogate *s, INT epochs, REAL lr, INT seed, int flags, REAL *loss_history) {
  const int augment = flags & GRID_SURROGATE_AUGMENT;
  grid_timer tm;
  grid_timer_start(&tm);
  if (!s->set_n) lg_fatal("grid_surrogate_train: no training data (grid_surrogate_set_training_data)");
  long *perm = (long *)malloc(sizeof(long) * s->set_n);
  unsigned long long st = 0xd1b54a32d192ed03ULL * (unsigned long long)(seed + 7);
  const long nbatches = (s->set_n + s->batch - 1) / s->batch;
  const long total = nbatches * epochs;
  /* linear warm-up: Adam's first steps move every weight by about lr at once */
  const long warmup = total / 20 > 20 ? total / 20 : 20;
  double last = 0.0;
  long it = 0;
  for (INT ep = 0; ep < epochs; ++ep) {
    for (long i = 0; i < s->set_n; ++i) perm[i] = i;
    for (long i = s->set_n - 1; i > 0; --i) {
      const long j = (long)(lg_splitmix64(&st) % (unsigned long long)(i + 1));
      const long t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
    double sum = 0.0;
    for (long bi = 0; bi < nbatches; ++bi, ++it) {
      const long i0 = bi * s->batch;
      const long nb = (s->set_n - i0) < s->batch ? (s->set_n - i0) : s->batch;
      for (long k = 0; k < nb; ++k) {
        const float *sv = s->set_v + perm[i0 + k] * s->nv, *stg = s->set_t + perm[i0 + k] * s->nv;
        if (augment) {
          /* the same random symmetry of the periodic cube for V and its ground state */
          const unsigned long long r = lg_splitmix64(&st);
          const int S = 1 << s->lgS, ax = (int)(r % 6), fl = (int)((r >> 8) & 7);
          const int s0 = (int)((r >> 16) % S), s1 = (int)((r >> 32) % S), s2 = (int)((r >> 48) % S);
          lg_hip(hgrid_nn_symgather(sv, s->vin + k * s->nv, s->lgS, ax, fl, s0, s1, s2), "nn_symgather");
          lg_hip(hgrid_nn_symgather(stg, s->tgt + k * s->nv, s->lgS, ax, fl, s0, s1, s2), "nn_symgather");
        } else {
          lg_hip(hgrid_memcpy_d2d(s->vin + k * s->nv, sv, sizeof(float) * s->nv), "batch");
          lg_hip(hgrid_memcpy_d2d(s->tgt + k * s->nv, stg, sizeof(float) * s->nv), "batch");
        }
      }
      forward(s, s->nv * nb);
      sum += backward(s, nb) * (double)nb;
      const double frac = (double)it / (double)(total > 1 ? total - 1 : 1);
      const double warm = (double)(it + 1) / (double)warmup;
      adam_step(s, lr * (warm < 1.0 ? warm : 1.0) * (0.05 + 0.95 * 0.5 * (1.0 + cos(M_PI * frac))));
    }
    last = sum / (double)s->set_n;
    if (loss_history) loss_history[ep] = last;
    if (flags & GRID_SURROGATE_PROGRESS) {
      float gn[2];
      lg_hip(hgrid_memcpy_d2h(gn, s->scal + 2, sizeof(gn)), "hipMemcpy (gradient norm)");
      fprintf(stderr, "libgrid: surrogate epoch %ld/%ld loss %.6f, last |g| %.3g (%.1f s)\n", (long)ep + 1, (long)epochs,
              last, (double)gn[1], grid_timer_wall_clock_time(&tm));
      fflush(stderr);
    }
  }
  free(perm);
  return last;
}

/* Mean loss of the current network over n samples (no update). */
REAL grid_surrogate_loss(grid_surrogate *s, INT n, const REAL *pot, const REAL *target) {
  double sum = 0.0;
  for (INT i0 = 0; i0 < n; i0 += s->batch) {
    const long nb = (n - i0) < s->batch ? (n - i0) : s->batch;
    upload_batch(s, nb, pot + i0 * s->nv, target + i0 * s->nv);
    forward(s, s->nv * nb);
    sum += backward(s, nb) * (double)nb;
  }
  return sum / (double)n;
}

/* ---- persistence ----------------------------------------------------------------------------- */

/* file: 8-byte magic, side, parameter count, 0 (int64), vscale (float64), FP32 parameters */
static const char kMagic[8] = {'L', 'G', 'U', 'N', 'E', 'T', '0', '2'};

int grid_surrogate_save(grid_surrogate *s, const char *path) {
  FILE *fp = fopen(path, "wb");
  if (!fp) return -1;
  float *h = (float *)malloc(sizeof(float) * s->np);
  lg_hip(hgrid_memcpy_d2h(h, s->w, sizeof(float) * s->np), "hipMemcpy (params)");
  const long hdr[3] = {1L << s->lgS, s->np, 0};
  const double vs = s->vscale;
  int ok = fwrite(kMagic, 1, 8, fp) == 8 && fwrite(hdr, sizeof(long), 3, fp) == 3 && fwrite(&vs, sizeof(double), 1, fp) == 1 &&
           fwrite(h, sizeof(float), s->np, fp) == (size_t)s->np;
  free(h);
  return (fclose(fp) == 0 && ok) ? 0 : -1;
}

grid_surrogate *grid_surrogate_load(const char *path, INT batch) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return NULL;
  char magic[8];
  long hdr[3];
  double vs;
  if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, kMagic, 8) || fread(hdr, sizeof(long), 3, fp) != 3 ||
      fread(&vs, sizeof(double), 1, fp) != 1) {
    fclose(fp);
    return NULL;
  }
  grid_surrogate *s = grid_surrogate_alloc(hdr[0], batch, vs, 0);
  if (s->np != hdr[1]) {
    fclose(fp);
    grid_surrogate_free(s);
    return NULL;
  }
  float *h = (float *)malloc(sizeof(float) * s->np);
  const size_t got = fread(h, sizeof(float), s->np, fp);
  fclose(fp);
  if (got != (size_t)s->np) {
    free(h);
    grid_surrogate_free(s);
    return NULL;
  }
  lg_hip(hgrid_memcpy_h2d(s->w, h, sizeof(float) * s->np), "hipMemcpy (params)");
  free(h);
  refresh_weights(s);
  return s;
}

/* ---- warm start ------------------------------------------------------------------------------ */

/* Plain start: exp(-r^2 / (2 width^2)) at the grid centre (x = y = z = 0), normalised. */
void grid_wf_gaussian_start(wf *gwf, REAL width) {
  cgrid *g = gwf->grid;
  const INT nx = g->nx, ny = g->ny, nz = g->nz;
  const double st = g->step, a = 1.0 / (2.0 * width * width);
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < nx; ++i) {
    const double x = (double)(i - nx / 2) * st;
    for (INT j = 0; j < ny; ++j) {
      const double y = (double)(j - ny / 2) * st;
      for (INT k = 0; k < nz; ++k) {
        const double z = (double)(k - nz / 2) * st;
        g->value[(i * ny + j) * nz + k] = exp(-a * (x * x + y * y + z * z));
      }
    }
  }
  cgrid_host_modified(g);
  grid_wf_normalize(gwf);
}

/* Network start: V sampled every n / S points (the S^3 subgrid shares the wavefunction grid's
 * origin), network amplitude, periodic trilinear interpolation back to n^3, normalised.  With
 * s == NULL the plain Gaussian of width 'fallback_width' is used. */
void grid_wf_warm_start(wf *gwf, rgrid *potential, grid_surrogate *s, REAL fallback_width) {
  if (!s) {
    grid_wf_gaussian_start(gwf, fallback_width);
    return;
  }
  cgrid *g = gwf->grid;
  const long S = 1L << s->lgS, n = g->nx;
  if (g->ny != n || g->nz != n || n % S || potential->nx != n || potential->ny != n || potential->nz != n)
    lg_fatal("grid_wf_warm_start: cubic n^3 grids with n a multiple of %ld", S);
  const long f = n / S;
  rgrid_host_sync(potential);
  double *v = (double *)malloc(sizeof(double) * s->nv);
  for (long i = 0; i < S; ++i)
    for (long j = 0; j < S; ++j)
      for (long k = 0; k < S; ++k)
        v[(i * S + j) * S + k] = potential->value[((i * f) * n + j * f) * potential->nz2 + k * f];
  upload_batch(s, 1, v, NULL);
  free(v);
  forward(s, s->nv);
  void *psi = lg_c_out(g);
  lg_hip(hgrid_nn_to_grid(s->out, s->lgS, (int)f, n, psi, g->dev->prec), "nn_to_grid");
  lg_c_written(g, 1);
  grid_wf_normalize(gwf);
}
