// Pointwise kernels and deterministic reductions behind the C-ABI (grid_hip.h).
//
// All kernels are grid-stride loops over flat element indices.  Reductions use a fixed
// launch geometry that depends only on n, each thread sums a fixed subset, the workgroup
// tree is fixed, and the final pass sums the per-workgroup partials in index order with a
// Neumaier-compensated FP64 accumulator, so results are bitwise reproducible run to run.
// Reference operators: cgrid_multiply (src/cgrid.c:1223-1238), grid_wf_propagate_potential
// (src/grid_wf.c:569-600), grid_wf_propagate_kinetic_fft (src/grid_wf_fft.c:141-193),
// cgrid_integral_of_square (src/cgrid.c:1610-1629), grid_grid_expectation_value
// (src/ngrid.c:355-376), cgrid_integral_of_conjugate_product (src/cgrid.c:1640-1658),
// grid_wf_density (src/grid_wf.c:611-627), grid_real_to_complex_re (src/ngrid.c:53-80),
// cgrid_fd_laplace (src/cgrid_diff.c:303-330) with the boundary rules of src/cgrid_bc.c:21-86.

#include "cx.hpp"
#include "grid_hip.h"
#include "phase_math.hpp"
#include "rt.hpp"

namespace lg {

constexpr int kNT = 256;

static inline unsigned grid_for(long n, long per_thread = 4) {
  long b = (n + (long)kNT * per_thread - 1) / ((long)kNT * per_thread);
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (unsigned)b;
}

#define LG_GRID_STRIDE(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

template <typename C> __device__ __forceinline__ C scal(const double *s) {
  using R = typename real_of<C>::type;
  return s ? mk<C>((R)s[0], (R)s[1]) : mk<C>(1, 0);
}

// ---- conversions -------------------------------------------------------------------------

__global__ void k_c128_to_c64(float2 *d, const double2 *s, long n) {
  LG_GRID_STRIDE(i, n) { d[i] = make_float2((float)s[i].x, (float)s[i].y); }
}
__global__ void k_c64_to_c128(double2 *d, const float2 *s, long n) {
  LG_GRID_STRIDE(i, n) { d[i] = make_double2((double)s[i].x, (double)s[i].y); }
}
__global__ void k_rpack(double *d, const double *s, long rows, long nz, long nz2) {
  const long n = rows * nz;
  LG_GRID_STRIDE(i, n) {
    const long r = i / nz, k = i - r * nz;
    d[i] = s[r * nz2 + k];
  }
}
__global__ void k_runpack(double *d, const double *s, long rows, long nz, long nz2) {
  const long n = rows * nz2;
  LG_GRID_STRIDE(i, n) {
    const long r = i / nz2, k = i - r * nz2;
    d[i] = k < nz ? s[r * nz + k] : 0.0;
  }
}
// One 16-byte element per thread and one thread per element: of the copy shapes measured on
// MI355X (grid-stride loops, 1-16 elements per thread, non-temporal loads and stores) this one
// moves the most, 6.1-6.2 TB/s on 2-4 GiB buffers against 4.5-5.2 TB/s for a grid-stride loop.
__global__ void __launch_bounds__(kNT) k_stream_copy(double2 *d, const double2 *s, long n) {
  const long i = (long)blockIdx.x * kNT + threadIdx.x;
  if (i < n) d[i] = s[i];
}

__global__ void k_fill_bytes(unsigned char *p, unsigned char v, long n) {
  LG_GRID_STRIDE(i, n) { p[i] = v; }
}

// ---- reductions --------------------------------------------------------------------------

// Deterministic block sum of NCH values per thread into out[c * gridDim.x + blockIdx.x].
template <int NCH> __device__ __forceinline__ void block_sum_store(double (&acc)[NCH], double *out) {
  __shared__ double sh[kNT / 64][NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    double x = acc[c];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    acc[c] = x;
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) sh[wave][c] = acc[c];
  }
  __syncthreads();
  if (threadIdx.x < NCH) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kNT / 64; ++w) s += sh[w][threadIdx.x];
    out[(long)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// One workgroup of 1024 threads sums one channel: each thread a fixed strided subset with
// Neumaier compensation, then a fixed-order tree; bitwise reproducible.
constexpr int kRT = 1024;
__device__ double reduce_channel(const double *x, long nb, double *sh_s, double *sh_c) {
  double s = 0.0, comp = 0.0;
  for (long b = threadIdx.x; b < nb; b += kRT) {
    const double v = x[b];
    const double t = s + v;
    if (fabs(s) >= fabs(v)) comp += (s - t) + v;
    else comp += (v - t) + s;
    s = t;
  }
  sh_s[threadIdx.x] = s;
  sh_c[threadIdx.x] = comp;
  __syncthreads();
  for (int w = kRT / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double a = sh_s[threadIdx.x], v = sh_s[threadIdx.x + w];
      const double t = a + v;
      const double e = fabs(a) >= fabs(v) ? (a - t) + v : (v - t) + a;
      sh_s[threadIdx.x] = t;
      sh_c[threadIdx.x] += sh_c[threadIdx.x + w] + e;
    }
    __syncthreads();
  }
  const double r = sh_s[0] + sh_c[0];
  __syncthreads();
  return r;
}

// One workgroup per reduced channel (blockIdx.x enumerates the channels of all jobs in
// order).  The workgroup that owns the channel holding the norm sum also writes the scale.
__global__ void __launch_bounds__(kRT) k_step_finalize(const hgrid_finalize f) {
  __shared__ double sh_s[kRT], sh_c[kRT];
  int b = blockIdx.x;
  for (int j = 0; j < f.njobs; ++j) {
    const hgrid_reduce_job jb = f.job[j];
    if (b < jb.nchan) {
      const double r = reduce_channel(jb.partials + (long)(jb.chan_first + b) * jb.nblocks, jb.nblocks, sh_s, sh_c);
      if (threadIdx.x == 0) {
        jb.out[b] = r;
        if (f.scale && f.target > 0.0 && jb.out + b == f.norm_sum) {
          f.scale[0] = sqrt(f.target / (r * f.dv));
          f.scale[1] = 0.0;
        }
      }
      return;
    }
    b -= jb.nchan;
  }
  // the extra workgroup (no channel left): reset a consumed pending scale
  if (threadIdx.x == 0 && f.scale && !(f.target > 0.0)) {
    f.scale[0] = 1.0;
    f.scale[1] = 0.0;
  }
}

template <typename C> __global__ void __launch_bounds__(kNT) k_sum_norm2(double *part, const C *a, long n, const double *scale) {
  double acc[1] = {0.0};
  LG_GRID_STRIDE(i, n) { acc[0] += cnorm2(a[i]); }
  if (scale) acc[0] *= scale[0] * scale[0] + scale[1] * scale[1];
  block_sum_store<1>(acc, part);
}

template <typename C>
__global__ void __launch_bounds__(kNT) k_sum_vnorm2(double *part, const C *a, const double *v, long n, const double *scale) {
  double acc[1] = {0.0};
  LG_GRID_STRIDE(i, n) { acc[0] += v[i] * cnorm2(a[i]); }
  if (scale) acc[0] *= scale[0] * scale[0] + scale[1] * scale[1];
  block_sum_store<1>(acc, part);
}

template <typename C> __global__ void __launch_bounds__(kNT) k_sum_conj_product(double *part, const C *a, const C *b, long n) {
  double acc[2] = {0.0, 0.0};
  LG_GRID_STRIDE(i, n) {
    const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
    acc[0] += ax * bx + ay * by;
    acc[1] += ax * by - ay * bx;
  }
  block_sum_store<2>(acc, part);
}

template <typename C>
__global__ void __launch_bounds__(kNT) k_sum_k2norm2(double *part, const C *a, long nx, long ny, long nz, const double *k2x,
                                                     const double *k2y, const double *k2z, const double *scale) {
  double acc[1] = {0.0};
  const long n = nx * ny * nz;
  LG_GRID_STRIDE(idx, n) {
    const long i = idx / (ny * nz), j = (idx / nz) % ny, k = idx % nz;
    acc[0] += (k2x[i] + k2y[j] + k2z[k]) * cnorm2(a[idx]);
  }
  if (scale) acc[0] *= scale[0] * scale[0] + scale[1] * scale[1];
  block_sum_store<1>(acc, part);
}

__global__ void __launch_bounds__(kNT) k_sum_real(double *part, const double *a, long n, int squared) {
  double acc[1] = {0.0};
  LG_GRID_STRIDE(i, n) { acc[0] += squared ? a[i] * a[i] : a[i]; }
  block_sum_store<1>(acc, part);
}

// sum Re(conj(psi) * lap7(psi)) * (1/h^2 folded by the caller); bc 2 periodic, 1 Neumann, 0 Dirichlet.
// The pending scale s multiplies the stored values only: the Dirichlet value stands for the
// boundary function's return value, which the scale does not touch.
template <typename C>
__device__ __forceinline__ double2 fd_at(const C *p, long nx, long ny, long nz, long i, long j, long k, int bc,
                                         double2 outside, double2 s) {
  if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) {
    if (bc == 2) {
      i %= nx; if (i < 0) i += nx;
      j %= ny; if (j < 0) j += ny;
      k %= nz; if (k < 0) k += nz;
    } else if (bc == 1) {
      if (i < 0) i = -i;
      if (i > nx - 1) i = 2 * nx - i - 2;
      if (j < 0) j = -j;
      if (j > ny - 1) j = 2 * ny - j - 2;
      if (k < 0) k = -k;
      if (k > nz - 1) k = 2 * nz - k - 2;
      if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return outside;
    } else {
      return outside;
    }
  }
  const C v = p[(i * ny + j) * nz + k];
  return make_double2(s.x * (double)v.x - s.y * (double)v.y, s.x * (double)v.y + s.y * (double)v.x);
}

template <typename C>
__global__ void __launch_bounds__(kNT) k_fd_laplace_expect(double *part, const C *p, long nx, long ny, long nz, int bc,
                                                           double2 outside, const double *scale) {
  double acc[1] = {0.0};
  const long n = nx * ny * nz;
  const double2 s = scale ? make_double2(scale[0], scale[1]) : make_double2(1.0, 0.0);
  LG_GRID_STRIDE(idx, n) {
    const long i = idx / (ny * nz), j = (idx / nz) % ny, k = idx % nz;
    const double2 c = fd_at(p, nx, ny, nz, i, j, k, bc, outside, s);
    const double2 nb[6] = {fd_at(p, nx, ny, nz, i, j, k + 1, bc, outside, s), fd_at(p, nx, ny, nz, i, j, k - 1, bc, outside, s),
                           fd_at(p, nx, ny, nz, i, j + 1, k, bc, outside, s), fd_at(p, nx, ny, nz, i, j - 1, k, bc, outside, s),
                           fd_at(p, nx, ny, nz, i + 1, j, k, bc, outside, s), fd_at(p, nx, ny, nz, i - 1, j, k, bc, outside, s)};
    double lr = -6.0 * c.x, li = -6.0 * c.y;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      lr += nb[m].x;
      li += nb[m].y;
    }
    acc[0] += c.x * lr + c.y * li;
  }
  block_sum_store<1>(acc, part);
}

// ---- pointwise ---------------------------------------------------------------------------

template <typename C> __device__ __forceinline__ C pfac(double cr, double ci, double vr, double vi);
template <> __device__ __forceinline__ double2 pfac<double2>(double cr, double ci, double vr, double vi) {
  const double ar = cr * vr - ci * vi, ai = cr * vi + ci * vr;
  double s, c;
  sincos_f64(ai, &s, &c);
  const double m = exp_f64(ar);
  return make_double2(m * c, m * s);
}
template <> __device__ __forceinline__ float2 pfac<float2>(double cr, double ci, double vr, double vi) {
  const double ar = cr * vr - ci * vi, ai = cr * vi + ci * vr;
  float s, c;
  sincos_f32(ai, &s, &c);
  const float m = expf((float)ar);
  return make_float2(m * c, m * s);
}

template <typename C>
__global__ void k_potential_phase(C *psi, const C *pot, long n, double cr, double ci, double cons, const double *scale) {
  const C s = scal<C>(scale);
  LG_GRID_STRIDE(i, n) {
    const C v = pot[i];
    C x = psi[i];
    if (scale) x = cmul(x, s);
    psi[i] = cmul(x, pfac<C>(cr, ci, cons + (double)v.x, (double)v.y));
  }
}

template <typename C>
__global__ void k_kinetic_phase(C *psi, long nx, long ny, long nz, const double2 *ta, const double2 *tb,
                                const double2 *tc, const double *scale, const double *kcx, const double *kcy,
                                const double *kcz, double kcut2) {
  const long n = nx * ny * nz;
  LG_GRID_STRIDE(idx, n) {
    const long i = idx / (ny * nz), j = (idx / nz) % ny, k = idx % nz;
    // FP64 tables (and pending scale) multiplied in FP64, rounded once on the store
    double2 f = cmul(ta[i], cmul(tb[j], tc[k]));
    if (scale) f = cmul(f, make_double2(scale[0], scale[1]));
    // CFFT cutoff, |k - k0|^2 summed in the order x, y, z; cut components become exact zeros
    if (kcx && (kcx[i] + kcy[j]) + kcz[k] >= kcut2) psi[idx] = mk<C>(0, 0);
    else psi[idx] = cmul_d(psi[idx], f);
  }
}

// absorbing-boundary amplitude: 0 inside [lx, hx] x [ly, hy] x [lz, hz]; outside, the distance
// over the lower bound per axis (both sides), summed, times 2/3, capped at 1
__device__ static double absorb_amp(long i, long j, long k, long lx, long hx, long ly, long hy, long lz, long hz) {
  double t = 0.0;
  if (i < lx) t += (double)(lx - i) / (double)lx;
  else if (i > hx) t += (double)(i - hx) / (double)lx;
  if (j < ly) t += (double)(ly - j) / (double)ly;
  else if (j > hy) t += (double)(j - hy) / (double)ly;
  if (k < lz) t += (double)(lz - k) / (double)lz;
  else if (k > hz) t += (double)(k - hz) / (double)lz;
  t *= 2.0 / 3.0;
  return t > 1.0 ? 1.0 : t;
}

// dst = (1 - alpha) wr + alpha wi (dst may be wr); products and sums rounded separately
template <typename C>
__global__ void k_wf_merge(C *dst, const C *wr, const C *wi, long nx, long ny, long nz, long lx, long hx, long ly,
                           long hy, long lz, long hz) {
#pragma clang fp contract(off)
  const long n = nx * ny * nz;
  LG_GRID_STRIDE(idx, n) {
    const long i = idx / (ny * nz), j = (idx / nz) % ny, k = idx % nz;
    const double al = absorb_amp(i, j, k, lx, hx, ly, hy, lz, hz), bl = 1.0 - al;
    const C r = wr[idx], m = wi[idx];
    const double re = bl * (double)r.x + al * (double)m.x, im = bl * (double)r.y + al * (double)m.y;
    dst[idx] = mk<C>((typename real_of<C>::type)re, (typename real_of<C>::type)im);
  }
}

// c += d a (complex d formed in FP64, product and sum rounded once to the storage precision)
template <typename C> __global__ void k_caxpy(C *c, const C *a, long n, double dr, double di) {
  LG_GRID_STRIDE(i, n) {
    const C x = a[i], y = c[i];
    const double re = (double)y.x + (dr * (double)x.x - di * (double)x.y);
    const double im = (double)y.y + (dr * (double)x.y + di * (double)x.x);
    c[i] = mk<C>((typename real_of<C>::type)re, (typename real_of<C>::type)im);
  }
}

template <typename C> __global__ void k_cscale(C *psi, long n, double cr, double ci, const double *scale) {
  C f = mk<C>((typename real_of<C>::type)cr, (typename real_of<C>::type)ci);
  if (scale) {
    // product formed in FP64 before rounding to the storage precision
    const double re = cr * scale[0] - ci * scale[1], im = cr * scale[1] + ci * scale[0];
    f = mk<C>((typename real_of<C>::type)re, (typename real_of<C>::type)im);
  }
  LG_GRID_STRIDE(i, n) { psi[i] = cmul(psi[i], f); }
}

template <typename C> __global__ void k_density(double *d, const C *s, long n, const double *scale) {
  const double f = scale ? scale[0] * scale[0] + scale[1] * scale[1] : 1.0;
  LG_GRID_STRIDE(i, n) { d[i] = cnorm2(s[i]) * f; }
}

template <typename C> __global__ void k_real_to_complex(C *d, const double *s, long n) {
  LG_GRID_STRIDE(i, n) { d[i] = mk<C>((typename real_of<C>::type)s[i], 0); }
}

template <typename C> __global__ void k_cfill(C *d, long n, double cr, double ci) {
  const C v = mk<C>((typename real_of<C>::type)cr, (typename real_of<C>::type)ci);
  LG_GRID_STRIDE(i, n) { d[i] = v; }
}

__global__ void k_scalar_set(double *s, double re, double im) {
  if (threadIdx.x == 0) {
    s[0] = re;
    s[1] = im;
  }
}
__global__ void k_scalar_mul(double *s, double re, double im) {
  if (threadIdx.x == 0) {
    const double a = s[0], b = s[1];
    s[0] = a * re - b * im;
    s[1] = a * im + b * re;
  }
}
__global__ void k_scalar_renorm(double *s, const double *sum, double target, double dv) {
  if (threadIdx.x == 0) {
    s[0] = sqrt(target / (sum[0] * dv));
    s[1] = 0.0;
  }
}

}  // namespace lg

using namespace lg;

#define LG_LAUNCH(kern, n, ...)                                                          \
  do {                                                                                   \
    hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kNT), 0, stream(), __VA_ARGS__);   \
    return check_launch(#kern);                                                          \
  } while (0)

extern "C" int hgrid_c128_to_c64(void *d, const void *s, size_t n) {
  LG_LAUNCH(k_c128_to_c64, (long)n, (float2 *)d, (const double2 *)s, (long)n);
}
extern "C" int hgrid_c64_to_c128(void *d, const void *s, size_t n) {
  LG_LAUNCH(k_c64_to_c128, (long)n, (double2 *)d, (const float2 *)s, (long)n);
}
extern "C" int hgrid_rpack(double *d, const double *s, long rows, long nz, long nz2) {
  LG_LAUNCH(k_rpack, rows * nz, d, s, rows, nz, nz2);
}
extern "C" int hgrid_runpack(double *d, const double *s, long rows, long nz, long nz2) {
  LG_LAUNCH(k_runpack, rows * nz2, d, s, rows, nz, nz2);
}
extern "C" int hgrid_stream_copy(void *d, const void *s, size_t bytes) {
  const long n = (long)(bytes / 16);
  const long nb = (n + kNT - 1) / kNT;
  if (n <= 0) return 0;
  if (nb > 0x7fffffffL) return fail("hgrid_stream_copy: buffer too large for one launch");
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)nb), dim3(kNT), 0, stream(), (double2 *)d, (const double2 *)s, n);
  return check_launch("stream_copy");
}

__global__ void k_copy_f64(double *d, const double *s, long n) {
  LG_GRID_STRIDE(i, n) d[i] = s[i];
}

// a kernel rather than hipMemcpyAsync, so that captured graphs hold kernel nodes only
extern "C" int hgrid_copy_f64(double *dst, const double *src, long n) { LG_LAUNCH(k_copy_f64, n, dst, src, n); }

extern "C" int hgrid_memset(void *p, int v, size_t bytes) {
  LG_LAUNCH(k_fill_bytes, (long)bytes, (unsigned char *)p, (unsigned char)v, (long)bytes);
}

extern "C" int hgrid_step_finalize(const hgrid_finalize *f) {
  if (f->njobs < 0 || f->njobs > HGRID_MAX_JOBS) return -1;
  int nb = 0;
  for (int j = 0; j < f->njobs; ++j) nb += f->job[j].nchan;
  if (f->scale && f->target > 0.0) {
    /* the norm sum must be one of the reduced outputs */
    int found = 0;
    for (int j = 0; j < f->njobs; ++j)
      for (int c = 0; c < f->job[j].nchan; ++c) found |= (f->job[j].out + c == f->norm_sum);
    if (!found) return -1;
  }
  if (f->scale && !(f->target > 0.0)) nb += 1;
  if (nb == 0) return 0;
  hipLaunchKernelGGL(k_step_finalize, dim3(nb), dim3(kRT), 0, stream(), *f);
  return check_launch("step_finalize");
}

extern "C" int hgrid_reduce_partials(double *out, const double *partials, long nb, int nchan) {
  hgrid_finalize f = {};
  f.njobs = 1;
  f.job[0].partials = partials;
  f.job[0].nblocks = nb;
  f.job[0].chan_first = 0;
  f.job[0].nchan = nchan;
  f.job[0].out = out;
  return hgrid_step_finalize(&f);
}

extern "C" long hgrid_reduce_scratch(long n) { return 2L * (long)grid_for(n) + 2; }


extern "C" int hgrid_sum_norm2(double *out, const void *a, int prec, long n, const double *scale, double *scratch) {
  const unsigned nb = grid_for(n);
  if (prec == HGRID_C128)
    hipLaunchKernelGGL(k_sum_norm2<double2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const double2 *)a, n, scale);
  else
    hipLaunchKernelGGL(k_sum_norm2<float2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const float2 *)a, n, scale);
  if (check_launch("sum_norm2")) return -1;
  return hgrid_reduce_partials(out, scratch, nb, 1);
}

// ---- allocation-free norm reduction ------------------------------------------------------
// hgrid_sum_norm2 with its partials in a fixed device array instead of caller or library
// scratch: grid_for() caps a reduction at 8192 workgroups, so 2 * 8192 + 2 doubles (the size
// hgrid_reduce_scratch gives for the largest grids, 128 KiB) hold the partials of every n.
// Each device holds its own copy until the process ends; its address is looked up once per
// device at library set-up (hgrid_init), and host code uses only that address.
constexpr long kNoAllocScratch = 2L * 8192 + 2;
constexpr long kNoAllocMaxN = 1L << 34;  // elements: far beyond any grid that fits in HBM
static __device__ __align__(256) double scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static double *g_scratch_dev[kMaxDevices];

int lg::noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail("allocation-free scratch: device index out of range");
  if (g_scratch_dev[device]) return 0;
  void *p = nullptr;
  if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (allocation-free scratch)")) return -1;
  g_scratch_dev[device] = (double *)p;
  return 0;
}

extern "C" int hgrid_sum_norm2_noalloc(double *out, const void *a, int prec, long n, const double *scale) {
  if (n < 1 || n > kNoAllocMaxN) return fail("hgrid_sum_norm2_noalloc: n must be in 1 .. 2^34");
  static_assert(2L * 8192 + 2 <= kNoAllocScratch, "scratch holds the partials of the largest reduction");
  int dev = 0;
  if (check(hipGetDevice(&dev), "hipGetDevice")) return -1;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_dev[dev]) return fail("hgrid_sum_norm2_noalloc: call hgrid_init first");
  return hgrid_sum_norm2(out, a, prec, n, scale, g_scratch_dev[dev]);
}

extern "C" int hgrid_sum_vnorm2(double *out, const void *a, int prec, const double *v, long n, const double *scale,
                                double *scratch) {
  const unsigned nb = grid_for(n);
  if (prec == HGRID_C128)
    hipLaunchKernelGGL(k_sum_vnorm2<double2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const double2 *)a, v, n, scale);
  else
    hipLaunchKernelGGL(k_sum_vnorm2<float2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const float2 *)a, v, n, scale);
  if (check_launch("sum_vnorm2")) return -1;
  return hgrid_reduce_partials(out, scratch, nb, 1);
}

extern "C" int hgrid_sum_conj_product(double *out2, const void *a, const void *b, int prec, long n, double *scratch) {
  const unsigned nb = grid_for(n);
  if (prec == HGRID_C128)
    hipLaunchKernelGGL(k_sum_conj_product<double2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const double2 *)a,
                       (const double2 *)b, n);
  else
    hipLaunchKernelGGL(k_sum_conj_product<float2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const float2 *)a,
                       (const float2 *)b, n);
  if (check_launch("sum_conj_product")) return -1;
  return hgrid_reduce_partials(out2, scratch, nb, 2);
}

extern "C" int hgrid_sum_k2norm2(double *out, const void *a, int prec, long nx, long ny, long nz, const double *k2x,
                                 const double *k2y, const double *k2z, const double *scale, double *scratch) {
  const long n = nx * ny * nz;
  const unsigned nb = grid_for(n);
  if (prec == HGRID_C128)
    hipLaunchKernelGGL(k_sum_k2norm2<double2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const double2 *)a, nx, ny, nz,
                       k2x, k2y, k2z, scale);
  else
    hipLaunchKernelGGL(k_sum_k2norm2<float2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const float2 *)a, nx, ny, nz,
                       k2x, k2y, k2z, scale);
  if (check_launch("sum_k2norm2")) return -1;
  return hgrid_reduce_partials(out, scratch, nb, 1);
}

extern "C" int hgrid_sum_real(double *out, const double *a, long n, int squared, double *scratch) {
  const unsigned nb = grid_for(n);
  hipLaunchKernelGGL(k_sum_real, dim3(nb), dim3(kNT), 0, stream(), scratch, a, n, squared);
  if (check_launch("sum_real")) return -1;
  return hgrid_reduce_partials(out, scratch, nb, 1);
}

extern "C" int hgrid_fd_laplace_expect(double *out, const void *psi, int prec, long nx, long ny, long nz, int bc,
                                       double dir_re, double dir_im, const double *scale, double *scratch) {
  const long n = nx * ny * nz;
  const unsigned nb = grid_for(n);
  const double2 outside = make_double2(dir_re, dir_im);
  if (prec == HGRID_C128)
    hipLaunchKernelGGL(k_fd_laplace_expect<double2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const double2 *)psi,
                       nx, ny, nz, bc, outside, scale);
  else
    hipLaunchKernelGGL(k_fd_laplace_expect<float2>, dim3(nb), dim3(kNT), 0, stream(), scratch, (const float2 *)psi,
                       nx, ny, nz, bc, outside, scale);
  if (check_launch("fd_laplace_expect")) return -1;
  return hgrid_reduce_partials(out, scratch, nb, 1);
}

extern "C" int hgrid_potential_phase(void *psi, const void *pot, int prec, long n, double cr, double ci, double cons,
                                     const double *scale) {
  if (prec == HGRID_C128)
    LG_LAUNCH(k_potential_phase<double2>, n, (double2 *)psi, (const double2 *)pot, n, cr, ci, cons, scale);
  LG_LAUNCH(k_potential_phase<float2>, n, (float2 *)psi, (const float2 *)pot, n, cr, ci, cons, scale);
}

extern "C" int hgrid_kinetic_phase(void *psi, int prec, long nx, long ny, long nz, const void *ta, const void *tb,
                                   const void *tc, const double *scale, const double *kcx, const double *kcy,
                                   const double *kcz, double kcut2) {
  const long n = nx * ny * nz;
  if (kcx && (!kcy || !kcz)) return -1;
  if (prec == HGRID_C128)
    LG_LAUNCH(k_kinetic_phase<double2>, n, (double2 *)psi, nx, ny, nz, (const double2 *)ta, (const double2 *)tb,
              (const double2 *)tc, scale, kcx, kcy, kcz, kcut2);
  LG_LAUNCH(k_kinetic_phase<float2>, n, (float2 *)psi, nx, ny, nz, (const double2 *)ta, (const double2 *)tb,
            (const double2 *)tc, scale, kcx, kcy, kcz, kcut2);
}

extern "C" int hgrid_wf_merge(void *dst, const void *wr, const void *wi, int prec, long nx, long ny, long nz,
                              const long *bounds) {
  const long n = nx * ny * nz;
  if (prec == HGRID_C128)
    LG_LAUNCH(k_wf_merge<double2>, n, (double2 *)dst, (const double2 *)wr, (const double2 *)wi, nx, ny, nz, bounds[0],
              bounds[1], bounds[2], bounds[3], bounds[4], bounds[5]);
  LG_LAUNCH(k_wf_merge<float2>, n, (float2 *)dst, (const float2 *)wr, (const float2 *)wi, nx, ny, nz, bounds[0],
            bounds[1], bounds[2], bounds[3], bounds[4], bounds[5]);
}

extern "C" int hgrid_caxpy(void *c, const void *a, int prec, long n, double dr, double di) {
  if (prec == HGRID_C128) LG_LAUNCH(k_caxpy<double2>, n, (double2 *)c, (const double2 *)a, n, dr, di);
  LG_LAUNCH(k_caxpy<float2>, n, (float2 *)c, (const float2 *)a, n, dr, di);
}

extern "C" int hgrid_cscale(void *psi, int prec, long n, double cr, double ci, const double *scale) {
  if (prec == HGRID_C128) LG_LAUNCH(k_cscale<double2>, n, (double2 *)psi, n, cr, ci, scale);
  LG_LAUNCH(k_cscale<float2>, n, (float2 *)psi, n, cr, ci, scale);
}

extern "C" int hgrid_density(double *d, const void *s, int prec, long n, const double *scale) {
  if (prec == HGRID_C128) LG_LAUNCH(k_density<double2>, n, d, (const double2 *)s, n, scale);
  LG_LAUNCH(k_density<float2>, n, d, (const float2 *)s, n, scale);
}

extern "C" int hgrid_real_to_complex(void *d, int prec, const double *s, long n) {
  if (prec == HGRID_C128) LG_LAUNCH(k_real_to_complex<double2>, n, (double2 *)d, s, n);
  LG_LAUNCH(k_real_to_complex<float2>, n, (float2 *)d, s, n);
}

extern "C" int hgrid_cfill(void *d, int prec, long n, double cr, double ci) {
  if (prec == HGRID_C128) LG_LAUNCH(k_cfill<double2>, n, (double2 *)d, n, cr, ci);
  LG_LAUNCH(k_cfill<float2>, n, (float2 *)d, n, cr, ci);
}

extern "C" int hgrid_scalar_set(double *s, double re, double im) {
  hipLaunchKernelGGL(k_scalar_set, dim3(1), dim3(64), 0, stream(), s, re, im);
  return check_launch("scalar_set");
}
extern "C" int hgrid_scalar_mul(double *s, double re, double im) {
  hipLaunchKernelGGL(k_scalar_mul, dim3(1), dim3(64), 0, stream(), s, re, im);
  return check_launch("scalar_mul");
}
extern "C" int hgrid_scalar_renorm(double *s, const double *sum, double target, double dv) {
  hipLaunchKernelGGL(k_scalar_renorm, dim3(1), dim3(64), 0, stream(), s, sum, target, dv);
  return check_launch("scalar_renorm");
}
