// Kernels of the warm-start network: a 3-D U-Net over cubic grids of side S = 2^lgS,
// channels-last activations [voxel][channel] in BF16, FP32 accumulation on the gfx950
// matrix cores (v_mfma_f32_16x16x32_bf16), FP32 master weights and Adam state.
//
// Voxel index m = ((b S + z) S + y) S + x; 3x3x3 convolutions wrap periodically inside each
// sample (the solver's boundary condition); weights are [Cout][27][Cin] with tap
// t = 9 (dz+1) + 3 (dy+1) + (dx+1) and y[m] = sum_t sum_ci x[m + d_t][ci] w[co][t][ci].
//
// MFMA 16x16x32 operand layout (lane l, g = l / 16, r = l % 16): A[m = r][k = 8g + i],
// B[k = 8g + i][n = r] (8 bf16 per lane), D[m = 4g + i][n = r] (4 floats per lane).
// K runs over 8-channel chunks c = (tap, 8-channel block); lane group g takes chunk kc + g of
// each 4-chunk step, so A and B agree on k whatever the channel count.
#include <hip/hip_runtime.h>

#include "grid_hip.h"
#include "rt.hpp"

namespace lg {
namespace nn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 ld8(const bf16 *p) { return *(const bf16x8 *)p; }
__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (bf16)0.f;
  return z;
}
__device__ __forceinline__ float f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 b(float v) { return (bf16)v; }

// neighbour of voxel (bz, z, y, x) (bz = b << 3 lgS) at tap t
__device__ __forceinline__ long nbr(long bz, int z, int y, int x, int t, int lgS) {
  const int msk = (1 << lgS) - 1;
  const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  return bz + ((long)((z + dz) & msk) << (2 * lgS)) + ((long)((y + dy) & msk) << lgS) + ((x + dx) & msk);
}

// ---- convolution: implicit GEMM, M = voxels, N = output channels, K = 27 Cin ----------------
// Each wave computes MT x NTL tiles of 16 x 16; a workgroup is 4 waves over consecutive voxels.
// With reflected, transposed weights (k_wprep) the same kernel computes the data gradient.
template <int MT, int NTL>
__global__ void __launch_bounds__(256) k_conv3(const bf16 *__restrict__ x, int xs, int xo, int cin,
                                               const bf16 *__restrict__ w, const float *__restrict__ bias,
                                               bf16 *__restrict__ y, int ys, int yo, const bf16 *__restrict__ mask,
                                               int mks, int mko, int lgS, long nvox, int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const long m0 = ((long)blockIdx.x * 4 + wave) * (16 * MT);
  const int co0 = blockIdx.y * (16 * NTL);
  const int msk = (1 << lgS) - 1;
  const int nch = 27 * (cin >> 3);
  long vb[MT];
  int vz[MT], vy[MT], vx[MT];
  bool vok[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const long m = m0 + 16 * mt + r;
    vok[mt] = m < nvox;
    const long mm = vok[mt] ? m : 0;
    vx[mt] = (int)(mm & msk);
    vy[mt] = (int)((mm >> lgS) & msk);
    vz[mt] = (int)((mm >> (2 * lgS)) & msk);
    vb[mt] = (mm >> (3 * lgS)) << (3 * lgS);
  }
  f32x4 acc[MT][NTL];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int j = 0; j < NTL; ++j) acc[mt][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // chunk kc + g of every step: (tap, channel block cb)
  int tap = 0, cb = 8 * g;
  while (cb >= cin) {
    cb -= cin;
    ++tap;
  }
  const long wrow = 27L * cin;
#pragma unroll 1
  for (int kc = 0; kc < nch; kc += 4) {
    const bool cok = tap < 27;
    bf16x8 a[MT], bw[NTL];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      a[mt] = zero8();
      if (cok && vok[mt]) a[mt] = ld8(x + nbr(vb[mt], vz[mt], vy[mt], vx[mt], tap, lgS) * xs + xo + cb);
    }
#pragma unroll
    for (int j = 0; j < NTL; ++j) {
      bw[j] = zero8();
      if (cok) bw[j] = ld8(w + (long)(co0 + 16 * j + r) * wrow + tap * cin + cb);
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int j = 0; j < NTL; ++j) acc[mt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], bw[j], acc[mt][j], 0, 0, 0);
    cb += 32;
    while (cb >= cin) {
      cb -= cin;
      ++tap;
    }
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int j = 0; j < NTL; ++j) {
      const int co = co0 + 16 * j + r;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long m = m0 + 16 * mt + 4 * g + i;
        if (m < nvox) {
          float v = acc[mt][j][i] + bv;
          if (relu) v = fmaxf(v, 0.f);
          if (mask && !(f(mask[m * mks + mko + co]) > 0.f)) v = 0.f;
          y[m * ys + yo + co] = b(v);
        }
      }
    }
}

// ---- weight gradient: dW[co][t][ci] = sum_m dY[m][co] X[m + d_t][ci] ----------------------
// Grid (chunks, 27 taps, Cout / (16 CT)).  A workgroup stages 128 voxels of dY (16 CT channels)
// and of X at the tap's neighbours into LDS transposed to [channel][voxel], then each wave
// takes 32 of the voxels as the K of one MFMA per (co tile, ci tile).  Every wave writes its
// own partial sums (no atomics); k_reduce adds the partials in a fixed order.
constexpr int WG_V = 128;       // voxels staged per step
constexpr int WG_P = WG_V + 8;  // LDS row pitch (bf16): 16 rows of a b128 read hit 16 distinct bank groups
constexpr int WG_MAXT = 12;     // ci tiles (Cin <= 192)

template <int CT>
__global__ void __launch_bounds__(256) k_conv3_wgrad(const bf16 *__restrict__ dy, int dys, int dyo,
                                                     const bf16 *__restrict__ x, int xs, int xo, int cin, int lgS,
                                                     long nvox, long vc, int cout, float *__restrict__ partial,
                                                     float *__restrict__ bpartial) {
  __shared__ __align__(16) bf16 ys_[16 * CT * WG_P];
  __shared__ __align__(16) bf16 xs_[16 * WG_MAXT * WG_P];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int tap = blockIdx.y, co0 = blockIdx.z * 16 * CT;
  const int cpt = cin >> 3, ntile = (cin + 15) >> 4;
  const int msk = (1 << lgS) - 1;
  const long v0 = (long)blockIdx.x * vc;
  // rows cin .. 16 ntile - 1 stay zero (Cin = 8)
  for (int i = threadIdx.x; i < (16 * ntile - cin) * WG_P; i += 256) xs_[cin * WG_P + i] = b(0.f);
  f32x4 acc[CT][WG_MAXT];
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int j = 0; j < WG_MAXT; ++j) acc[c][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the centre tap's workgroups also sum dY over their voxels: the bias gradient
  const bool bias_wg = bpartial && tap == 13;
  float bsum = 0.f;
#pragma unroll 1
  for (long sb = v0; sb < v0 + vc; sb += WG_V) {
    __syncthreads();  // previous step's reads are done
    // dY: 128 voxels x 16 CT channels = 256 CT loads of 8 channels
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int idx = threadIdx.x + 256 * c;
      const int vl = idx & (WG_V - 1), h = (idx >> 7) * 8;  // h < 16 CT
      const long m = sb + vl;
      bf16x8 t = zero8();
      if (m < nvox) t = ld8(dy + m * dys + dyo + co0 + h);
#pragma unroll
      for (int i = 0; i < 8; ++i) ys_[(h + i) * WG_P + vl] = t[i];
    }
    // X at the tap's neighbours: 128 voxels x cpt chunks
    for (int idx = threadIdx.x; idx < WG_V * cpt; idx += 256) {
      const int vl = idx & (WG_V - 1), c8 = (idx >> 7) * 8;
      const long m = sb + vl;
      bf16x8 t = zero8();
      if (m < nvox) {
        const long bz = (m >> (3 * lgS)) << (3 * lgS);
        const long nb = nbr(bz, (int)((m >> (2 * lgS)) & msk), (int)((m >> lgS) & msk), (int)(m & msk), tap, lgS);
        t = ld8(x + nb * xs + xo + c8);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) xs_[(c8 + i) * WG_P + vl] = t[i];
    }
    __syncthreads();
    if (bias_wg && threadIdx.x < 16 * CT) {
      const bf16 *row = ys_ + threadIdx.x * WG_P;
      float s8 = 0.f;
      for (int i = 0; i < WG_V; ++i) s8 += f(row[i]);
      bsum += s8;
    }
    const int kv = 32 * wave + 8 * g;  // this lane's 8 voxels of the wave's 32
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const bf16x8 a = *(const bf16x8 *)(ys_ + (16 * c + r) * WG_P + kv);
#pragma unroll
      for (int j = 0; j < WG_MAXT; ++j)
        if (j < ntile) {
          const bf16x8 bb = *(const bf16x8 *)(xs_ + (16 * j + r) * WG_P + kv);
          acc[c][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[c][j], 0, 0, 0);
        }
    }
  }
  if (bias_wg && threadIdx.x < 16 * CT) bpartial[(long)blockIdx.x * cout + co0 + threadIdx.x] = bsum;
  // partial[(chunk * 4 + wave)][co][tap][ci]
  float *pp = partial + ((long)blockIdx.x * 4 + wave) * ((long)cout * 27 * cin);
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int j = 0; j < WG_MAXT; ++j)
      if (j < ntile) {
        const int ci = 16 * j + r;
        if (ci < cin) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int co = co0 + 16 * c + 4 * g + i;
            pp[((long)co * 27 + tap) * cin + ci] = acc[c][j][i];
          }
        }
      }
}

// out[e] = sum_p partial[p][e] in order p = 0, 1, ... (deterministic); accumulate adds to out
__global__ void k_reduce(const float *__restrict__ partial, long np, long n, float *__restrict__ out, int accumulate) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  // eight independent running sums (partials p = 8k + i go to sum i), added in a fixed order
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  long p = 0;
  for (; p + 8 <= np; p += 8)
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] += partial[(p + i) * n + e];
  for (int i = 0; p < np; ++p, ++i) s[i] += partial[p * n + e];
  const float t = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
  out[e] = accumulate ? out[e] + t : t;
}

// out[c] = sum_p partial[p * stride + c], c < ncols (fixed order)
// one workgroup per column: thread t takes rows t, t + 256, ...; then a fixed tree
__global__ void __launch_bounds__(256) k_reduce_strided(const float *__restrict__ partial, long np, int stride,
                                                        float *__restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (long p = threadIdx.x; p < np; p += 256) s += partial[p * stride + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// column sums of a [nvox][C] slice: partial[chunk][c]
__global__ void k_colsum(const bf16 *__restrict__ d, int ds, int dof, int C, long nvox, long vc, float *__restrict__ partial) {
  const long v0 = (long)blockIdx.x * vc;
  const long v1 = v0 + vc < nvox ? v0 + vc : nvox;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float s = 0.f;
    for (long m = v0; m < v1; ++m) s += f(d[m * ds + dof + c]);
    partial[(long)blockIdx.x * C + c] = s;
  }
}

// ---- resampling between levels (8-channel granules) -----------------------------------------
// coarse voxel of fine voxel m (fine side 2^lgF)
__device__ __forceinline__ long parent(long m, int lgF) {
  const int msk = (1 << lgF) - 1, lgC = lgF - 1;
  const long bb = m >> (3 * lgF);
  const int z = (int)((m >> (2 * lgF)) & msk) >> 1, y = (int)((m >> lgF) & msk) >> 1, x = (int)(m & msk) >> 1;
  return (bb << (3 * lgC)) + ((long)z << (2 * lgC)) + ((long)y << lgC) + x;
}
// k-th child (k = 4 kz + 2 ky + kx) of coarse voxel m (coarse side 2^lgC)
__device__ __forceinline__ long child(long m, int k, int lgC) {
  const int msk = (1 << lgC) - 1, lgF = lgC + 1;
  const long bb = m >> (3 * lgC);
  const int z = 2 * (int)((m >> (2 * lgC)) & msk) + (k >> 2), y = 2 * (int)((m >> lgC) & msk) + ((k >> 1) & 1),
            x = 2 * (int)(m & msk) + (k & 1);
  return (bb << (3 * lgF)) + ((long)z << (2 * lgF)) + ((long)y << lgF) + x;
}

// nearest upsampling: dst[m][c] = src[parent(m)][c]; grid over fine voxels x C / 8
__global__ void k_up2(const bf16 *__restrict__ src, int ss, int so, int C, bf16 *__restrict__ dst, int ds, int dof,
                      int lgF, long nfine) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cpt = C >> 3;
  if (id >= nfine * cpt) return;
  const long m = id / cpt;
  const int c8 = (int)(id % cpt) * 8;
  *(bf16x8 *)(dst + m * ds + dof + c8) = ld8(src + parent(m, lgF) * ss + so + c8);
}

// gradient of k_up2 times the ReLU mask of the coarse activation:
// dst[m][c] = (sum over the 8 children of dfine) * (act[m][c] > 0); grid over coarse voxels x C / 8
__global__ void k_up2_bwd(const bf16 *__restrict__ dfine, int fs, int fo, int C, const bf16 *__restrict__ act, int as,
                          int ao, bf16 *__restrict__ dst, int lgC, long ncoarse) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cpt = C >> 3;
  if (id >= ncoarse * cpt) return;
  const long m = id / cpt;
  const int c8 = (int)(id % cpt) * 8;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bf16x8 v = ld8(dfine + child(m, k, lgC) * fs + fo + c8);
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] += f(v[i]);
  }
  const bf16x8 a = ld8(act + m * as + ao + c8);
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; ++i) o[i] = b(f(a[i]) > 0.f ? s[i] : 0.f);
  *(bf16x8 *)(dst + m * C + c8) = o;
}

// 2x2x2 average pooling: dst[m][c] (dense, C channels) = mean of the 8 children; grid over coarse x C / 8
__global__ void k_pool2(const bf16 *__restrict__ src, int ss, int so, int C, bf16 *__restrict__ dst, int lgC,
                        long ncoarse) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cpt = C >> 3;
  if (id >= ncoarse * cpt) return;
  const long m = id / cpt;
  const int c8 = (int)(id % cpt) * 8;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bf16x8 v = ld8(src + child(m, k, lgC) * ss + so + c8);
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] += f(v[i]);
  }
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; ++i) o[i] = b(0.125f * s[i]);
  *(bf16x8 *)(dst + m * C + c8) = o;
}

// gradient into a pooled activation plus its skip gradient, times its ReLU mask:
// dst[m][c] = (dcoarse[parent(m)][c] / 8 + dskip[m][c]) * (act[m][c] > 0); grid over fine x C / 8
__global__ void k_pool2_bwd(const bf16 *__restrict__ dcoarse, const bf16 *__restrict__ dskip, int ks, int ko, int C,
                            const bf16 *__restrict__ act, int as, int ao, bf16 *__restrict__ dst, int lgF, long nfine) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cpt = C >> 3;
  if (id >= nfine * cpt) return;
  const long m = id / cpt;
  const int c8 = (int)(id % cpt) * 8;
  const bf16x8 dc = ld8(dcoarse + parent(m, lgF) * C + c8);
  const bf16x8 a = ld8(act + m * as + ao + c8);
  bf16x8 sk = zero8();
  if (dskip) sk = ld8(dskip + m * ks + ko + c8);
  bf16x8 o;
#pragma unroll
  for (int i = 0; i < 8; ++i) o[i] = b(f(a[i]) > 0.f ? 0.125f * f(dc[i]) + f(sk[i]) : 0.f);
  *(bf16x8 *)(dst + m * C + c8) = o;
}

// ---- 1x1 head and loss ------------------------------------------------------------------------
// out[m] = sum_c d[m][c] wh[c] + bh  (d: 16 channels)
__global__ void k_head_fwd(const bf16 *__restrict__ d, const float *__restrict__ wh, float *__restrict__ out, long nvox) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nvox) return;
  const bf16x8 a0 = ld8(d + m * 16), a1 = ld8(d + m * 16 + 8);
  float s = wh[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) s += f(a0[i]) * wh[i] + f(a1[i]) * wh[8 + i];
  out[m] = s;
}

// loss = scale/2 * sum (out - t)^2 with dout = scale (out - t); writes the ReLU-masked input
// gradient dpre[m][c] = dout wh[c] (d[m][c] > 0) and per-block partial sums
// partial[block][0..15] = sum dout d[.][c], [16] = sum dout, [17] = sum (out - t)^2.
__global__ void __launch_bounds__(256) k_head_bwd(const bf16 *__restrict__ d, const float *__restrict__ wh,
                                                  const float *__restrict__ out, const float *__restrict__ target,
                                                  float scale, bf16 *__restrict__ dpre, long nvox,
                                                  float *__restrict__ partial) {
  __shared__ float red[4][18];
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float v[18];
#pragma unroll
  for (int i = 0; i < 18; ++i) v[i] = 0.f;
  if (m < nvox) {
    const float e = out[m] - target[m];
    const float dout = scale * e;
    const bf16x8 a0 = ld8(d + m * 16), a1 = ld8(d + m * 16 + 8);
    bf16x8 o0, o1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      v[i] = dout * f(a0[i]);
      v[8 + i] = dout * f(a1[i]);
      o0[i] = b(f(a0[i]) > 0.f ? dout * wh[i] : 0.f);
      o1[i] = b(f(a1[i]) > 0.f ? dout * wh[8 + i] : 0.f);
    }
    v[16] = dout;
    v[17] = e * e;
    *(bf16x8 *)(dpre + m * 16) = o0;
    *(bf16x8 *)(dpre + m * 16 + 8) = o1;
  }
#pragma unroll
  for (int i = 0; i < 18; ++i)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < 18; ++i) red[wave][i] = v[i];
  __syncthreads();
  if (threadIdx.x < 18)
    partial[(long)blockIdx.x * 18 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// ---- parameters ---------------------------------------------------------------------------------
// BF16 copies of W[co][27][ci]: wb (same layout) and wt[ci][t][co] = W[co][26 - t][ci] (the
// data gradient as a forward convolution)
__global__ void k_wprep(const float *__restrict__ w, int cout, int cin, bf16 *__restrict__ wb, bf16 *__restrict__ wt) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = 27L * cout * cin;
  if (e >= n) return;
  const int ci = (int)(e % cin), t = (int)((e / cin) % 27), co = (int)(e / (27L * cin));
  const bf16 v = b(w[e]);
  wb[e] = v;
  wt[((long)ci * 27 + (26 - t)) * cout + co] = v;
}

// global gradient norm for clipping: per-block sums of squares (FP64), then one workgroup
// turns them into the factor min(1, clip / |g|) (0 when |g| is not finite: the step is skipped)
__global__ void __launch_bounds__(256) k_sumsq(const float *__restrict__ g, long n, double *__restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) s += (double)g[e] * g[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_clip_factor(const double *__restrict__ partial, int nb, double clip,
                                                     float *__restrict__ factor) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nrm = sqrt(red[0]);
    factor[0] = (float)(isfinite(nrm) ? (nrm > clip ? clip / nrm : 1.0) : 0.0);
    factor[1] = (float)nrm;
  }
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m1,
                       float *__restrict__ m2, long n, float lr, float b1, float b2, float eps, float c1, float c2,
                       const float *__restrict__ gscale) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float f = gscale ? gscale[0] : 1.f;
  if (f == 0.f) return;  // non-finite gradient: no update
  const float gg = gr[e] * f;
  const float a = b1 * m1[e] + (1.f - b1) * gg;
  const float s = b2 * m2[e] + (1.f - b2) * gg * gg;
  m1[e] = a;
  m2[e] = s;
  w[e] -= lr * (a * c1) / (sqrtf(s * c2) + eps);
}

// network input: channel 0 = vscale * V, channels 1..7 zero
__global__ void k_input(const float *__restrict__ v, float vscale, bf16 *__restrict__ x0, long nvox) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nvox) return;
  bf16x8 o = zero8();
  o[0] = b(vscale * v[m]);
  *(bf16x8 *)(x0 + m * 8) = o;
}

// dst = src under one symmetry of the periodic cube: dst[c] = src[c'] with the index triple
// c' = P(c) (axis permutation perm < 6), reflected i -> -i mod S on the axes in the bits of
// flips, then shifted by (s0, s1, s2) mod S.  The Schroedinger problem on the periodic box is
// invariant under these maps, so (V, psi_0) pairs stay pairs.
__global__ void k_symgather(const float *__restrict__ src, float *__restrict__ dst, int lgS, int perm, int flips, int s0,
                            int s1, int s2) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int S = 1 << lgS, msk = S - 1;
  if (m >= (1L << (3 * lgS))) return;
  const int c[3] = {(int)(m >> (2 * lgS)), (int)((m >> lgS) & msk), (int)(m & msk)};
  static constexpr int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int v = c[P[perm][a]];
    if ((flips >> a) & 1) v = (S - v) & msk;
    q[a] = v;
  }
  q[0] = (q[0] + s0) & msk;
  q[1] = (q[1] + s1) & msk;
  q[2] = (q[2] + s2) & msk;
  dst[m] = src[((long)q[0] << (2 * lgS)) + ((long)q[1] << lgS) + q[2]];
}

// periodic trilinear interpolation of a coarse S^3 field onto the n^3 grid (n = S * fct):
// fine index i sits at coarse coordinate i / fct.  psi gets the real value, zero imaginary part.
__global__ void k_to_grid(const float *__restrict__ a, int lgS, int fct, long n, void *__restrict__ psi, int prec) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n * n * n) return;
  const int S = 1 << lgS, msk = S - 1;
  const int i = (int)(id / (n * n)), j = (int)((id / n) % n), k = (int)(id % n);
  const int i0 = i / fct, j0 = j / fct, k0 = k / fct;
  const float ti = (float)(i - i0 * fct) / fct, tj = (float)(j - j0 * fct) / fct, tk = (float)(k - k0 * fct) / fct;
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
    const float wgt = (di ? ti : 1.f - ti) * (dj ? tj : 1.f - tj) * (dk ? tk : 1.f - tk);
    s += wgt * a[(((i0 + di) & msk) * S + ((j0 + dj) & msk)) * S + ((k0 + dk) & msk)];
  }
  if (prec == HGRID_C128) ((double2 *)psi)[id] = make_double2((double)s, 0.0);
  else ((float2 *)psi)[id] = make_float2(s, 0.f);
}

inline unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace nn
}  // namespace lg

using namespace lg::nn;

extern "C" int hgrid_nn_conv3(const void *x, int xs, int xo, int cin, const void *w, const float *bias, void *y, int ys,
                              int yo, int cout, const void *mask, int mks, int mko, int lgS, long nvox, int relu) {
  if (cin % 8 || cin < 8 || cout % 16 || xo + cin > xs || yo + cout > ys || (mask && mko + cout > mks) || lgS < 1 ||
      nvox <= 0 || nvox % (1L << (3 * lgS)))
    return lg::fail("nn_conv3: shape");
  if (((uintptr_t)x | (uintptr_t)w) & 15) return lg::fail("nn_conv3: operands must be 16-byte aligned");
  const int t16 = cout / 16;
  const int ntl = t16 % 4 == 0 ? 4 : t16 % 3 == 0 ? 3 : t16 % 2 == 0 ? 2 : 1;
  const int mt = ntl == 1 ? 4 : 2;
  dim3 grid(blocks(nvox, 64 * mt), t16 / ntl);
  auto *X = (const bf16 *)x;
  auto *W = (const bf16 *)w;
  auto *Y = (bf16 *)y;
  auto *M = (const bf16 *)mask;
  switch (ntl) {
    case 4: hipLaunchKernelGGL((k_conv3<2, 4>), grid, dim3(256), 0, lg::stream(), X, xs, xo, cin, W, bias, Y, ys, yo, M, mks, mko, lgS, nvox, relu); break;
    case 3: hipLaunchKernelGGL((k_conv3<2, 3>), grid, dim3(256), 0, lg::stream(), X, xs, xo, cin, W, bias, Y, ys, yo, M, mks, mko, lgS, nvox, relu); break;
    case 2: hipLaunchKernelGGL((k_conv3<2, 2>), grid, dim3(256), 0, lg::stream(), X, xs, xo, cin, W, bias, Y, ys, yo, M, mks, mko, lgS, nvox, relu); break;
    default: hipLaunchKernelGGL((k_conv3<4, 1>), grid, dim3(256), 0, lg::stream(), X, xs, xo, cin, W, bias, Y, ys, yo, M, mks, mko, lgS, nvox, relu); break;
  }
  return lg::check_launch("nn_conv3");
}

extern "C" long hgrid_nn_wgrad_chunk(long nvox) {
  // voxels per workgroup: a multiple of 128, at most 256 chunks
  long vc = (nvox + 255) / 256;
  vc = (vc + WG_V - 1) / WG_V * WG_V;
  return vc < WG_V ? WG_V : vc;
}

extern "C" long hgrid_nn_wgrad_partials(long nvox) {
  const long vc = hgrid_nn_wgrad_chunk(nvox);
  return 4 * ((nvox + vc - 1) / vc);
}

extern "C" long hgrid_nn_wgrad_chunks(long nvox) {
  const long vc = hgrid_nn_wgrad_chunk(nvox);
  return (nvox + vc - 1) / vc;
}

extern "C" int hgrid_nn_conv3_wgrad(const void *dy, int dys, int dyo, int cout, const void *x, int xs, int xo, int cin,
                                    int lgS, long nvox, float *partial, float *bpartial) {
  if (cin % 8 || cin > 16 * WG_MAXT || cout % 16 || xo + cin > xs || dyo + cout > dys || nvox <= 0)
    return lg::fail("nn_conv3_wgrad: shape");
  const long vc = hgrid_nn_wgrad_chunk(nvox);
  const int t16 = cout / 16;
  const int ct = t16 % 2 == 0 ? 2 : 1;  // LDS: 16 CT + 16 WG_MAXT rows of WG_P bf16 stay under 64 KiB
  dim3 grid((unsigned)((nvox + vc - 1) / vc), 27, t16 / ct);
  auto *D = (const bf16 *)dy;
  auto *X = (const bf16 *)x;
  switch (ct) {
    case 2: hipLaunchKernelGGL((k_conv3_wgrad<2>), grid, dim3(256), 0, lg::stream(), D, dys, dyo, X, xs, xo, cin, lgS, nvox, vc, cout, partial, bpartial); break;
    default: hipLaunchKernelGGL((k_conv3_wgrad<1>), grid, dim3(256), 0, lg::stream(), D, dys, dyo, X, xs, xo, cin, lgS, nvox, vc, cout, partial, bpartial); break;
  }
  return lg::check_launch("nn_conv3_wgrad");
}

extern "C" int hgrid_nn_reduce(const float *partial, long np, long n, float *out, int accumulate) {
  if (n <= 0 || np <= 0) return lg::fail("nn_reduce: shape");
  hipLaunchKernelGGL(k_reduce, dim3(blocks(n, 256)), dim3(256), 0, lg::stream(), partial, np, n, out, accumulate);
  return lg::check_launch("nn_reduce");
}

extern "C" int hgrid_nn_reduce_strided(const float *partial, long np, int stride, int ncols, float *out) {
  if (ncols <= 0 || stride < ncols || np <= 0) return lg::fail("nn_reduce_strided: shape");
  hipLaunchKernelGGL(k_reduce_strided, dim3(ncols), dim3(256), 0, lg::stream(), partial, np, stride, out);
  return lg::check_launch("nn_reduce_strided");
}

extern "C" long hgrid_nn_colsum_chunks(long nvox) { return (nvox + 1023) / 1024; }

extern "C" int hgrid_nn_colsum(const void *d, int ds, int dof, int c, long nvox, float *partial) {
  if (c <= 0 || dof + c > ds || nvox <= 0) return lg::fail("nn_colsum: shape");
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)hgrid_nn_colsum_chunks(nvox)), dim3(256), 0, lg::stream(), (const bf16 *)d,
                     ds, dof, c, nvox, 1024L, partial);
  return lg::check_launch("nn_colsum");
}

extern "C" int hgrid_nn_up2(const void *src, int ss, int so, int c, void *dst, int ds, int dof, int lgF, long nfine) {
  if (c % 8 || so + c > ss || dof + c > ds || lgF < 1) return lg::fail("nn_up2: shape");
  hipLaunchKernelGGL(k_up2, dim3(blocks(nfine * (c / 8), 256)), dim3(256), 0, lg::stream(), (const bf16 *)src, ss, so, c,
                     (bf16 *)dst, ds, dof, lgF, nfine);
  return lg::check_launch("nn_up2");
}

extern "C" int hgrid_nn_up2_bwd(const void *dfine, int fs, int fo, int c, const void *act, int as, int ao, void *dst,
                                int lgC, long ncoarse) {
  if (c % 8 || fo + c > fs || ao + c > as) return lg::fail("nn_up2_bwd: shape");
  hipLaunchKernelGGL(k_up2_bwd, dim3(blocks(ncoarse * (c / 8), 256)), dim3(256), 0, lg::stream(), (const bf16 *)dfine, fs,
                     fo, c, (const bf16 *)act, as, ao, (bf16 *)dst, lgC, ncoarse);
  return lg::check_launch("nn_up2_bwd");
}

extern "C" int hgrid_nn_pool2(const void *src, int ss, int so, int c, void *dst, int lgC, long ncoarse) {
  if (c % 8 || so + c > ss) return lg::fail("nn_pool2: shape");
  hipLaunchKernelGGL(k_pool2, dim3(blocks(ncoarse * (c / 8), 256)), dim3(256), 0, lg::stream(), (const bf16 *)src, ss, so,
                     c, (bf16 *)dst, lgC, ncoarse);
  return lg::check_launch("nn_pool2");
}

extern "C" int hgrid_nn_pool2_bwd(const void *dcoarse, const void *dskip, int ks, int ko, int c, const void *act, int as,
                                  int ao, void *dst, int lgF, long nfine) {
  if (c % 8 || (dskip && ko + c > ks) || ao + c > as || lgF < 1) return lg::fail("nn_pool2_bwd: shape");
  hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks(nfine * (c / 8), 256)), dim3(256), 0, lg::stream(), (const bf16 *)dcoarse,
                     (const bf16 *)dskip, ks, ko, c, (const bf16 *)act, as, ao, (bf16 *)dst, lgF, nfine);
  return lg::check_launch("nn_pool2_bwd");
}

extern "C" int hgrid_nn_head_fwd(const void *d, const float *wh, float *out, long nvox) {
  hipLaunchKernelGGL(k_head_fwd, dim3(blocks(nvox, 256)), dim3(256), 0, lg::stream(), (const bf16 *)d, wh, out, nvox);
  return lg::check_launch("nn_head_fwd");
}

extern "C" long hgrid_nn_head_blocks(long nvox) { return (nvox + 255) / 256; }

extern "C" int hgrid_nn_head_bwd(const void *d, const float *wh, const float *out, const float *target, float scale,
                                 void *dpre, long nvox, float *partial) {
  hipLaunchKernelGGL(k_head_bwd, dim3(blocks(nvox, 256)), dim3(256), 0, lg::stream(), (const bf16 *)d, wh, out, target,
                     scale, (bf16 *)dpre, nvox, partial);
  return lg::check_launch("nn_head_bwd");
}

extern "C" int hgrid_nn_wprep(const float *w, int cout, int cin, void *wb, void *wt) {
  const long n = 27L * cout * cin;
  hipLaunchKernelGGL(k_wprep, dim3(blocks(n, 256)), dim3(256), 0, lg::stream(), w, cout, cin, (bf16 *)wb, (bf16 *)wt);
  return lg::check_launch("nn_wprep");
}

extern "C" int hgrid_nn_clip_factor(const float *g, long n, double clip, double *scratch, float *factor) {
  const int nb = 512;
  hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, lg::stream(), g, n, scratch);
  hipLaunchKernelGGL(k_clip_factor, dim3(1), dim3(256), 0, lg::stream(), scratch, nb, clip, factor);
  return lg::check_launch("nn_clip_factor");
}

extern "C" int hgrid_nn_adam(float *w, const float *g, float *m1, float *m2, long n, float lr, float b1, float b2,
                             float eps, int t, const float *gscale) {
  const float c1 = 1.f / (1.f - powf(b1, (float)t)), c2 = 1.f / (1.f - powf(b2, (float)t));
  hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, lg::stream(), w, g, m1, m2, n, lr, b1, b2, eps, c1, c2,
                     gscale);
  return lg::check_launch("nn_adam");
}

extern "C" int hgrid_nn_input(const float *v, float vscale, void *x0, long nvox) {
  hipLaunchKernelGGL(k_input, dim3(blocks(nvox, 256)), dim3(256), 0, lg::stream(), v, vscale, (bf16 *)x0, nvox);
  return lg::check_launch("nn_input");
}

extern "C" int hgrid_nn_symgather(const float *src, float *dst, int lgS, int perm, int flips, int s0, int s1, int s2) {
  if (perm < 0 || perm > 5 || lgS < 1) return lg::fail("nn_symgather: arguments");
  hipLaunchKernelGGL(k_symgather, dim3(blocks(1L << (3 * lgS), 256)), dim3(256), 0, lg::stream(), src, dst, lgS, perm,
                     flips, s0, s1, s2);
  return lg::check_launch("nn_symgather");
}

extern "C" int hgrid_nn_to_grid(const float *a, int lgS, int fct, long n, void *psi, int prec) {
  if (fct < 1 || n != ((long)fct << lgS)) return lg::fail("nn_to_grid: shape");
  hipLaunchKernelGGL(k_to_grid, dim3(blocks(n * n * n, 256)), dim3(256), 0, lg::stream(), a, lgS, fct, n, psi, prec);
  return lg::check_launch("nn_to_grid");
}
