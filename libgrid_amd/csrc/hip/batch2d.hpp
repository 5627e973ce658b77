// Batched 2-D imaginary-time ground states in one persistent launch (complex64 storage,
// FP64 reductions).
//
// Each workgroup takes whole solves from an atomic work queue (the host orders the queue
// largest grid first) and relaxes one n x n grid at a time.  A grid of at most 256^2
// complex64 values (512 KiB) stays in L2 / Infinity Cache while its workgroup sweeps it.
//
// Pencils are owned by single waves: a wave holds 64 / TPP pencils, TPP lanes each with R
// values, and transforms them with an R-point DFT in registers, an exchange through the
// wave's own LDS region and TPP-point DFTs (the scheme of fft_pass.hpp).  No workgroup
// barrier sits inside a sweep, so the waves of a workgroup overlap each other's memory
// latency; the only barriers separate sweeps along different axes.
//
// Time stepping.  The reference step is psi_{n+1} = N[K V K psi_n] with K = exp(-(dt/2) k^2/2m)
// (k-space), V = exp(-dt V(r)) and N the renormalisation (grid_wf_propagate, src/grid_wf.c:379-432,
// in imaginary time, then grid_wf_normalize :694-700).  The kernel iterates the state between
// the two kinetic factors instead, chi_n with psi_n = N[K chi_n]:
//     chi_{n+1} = N[V K(dt) chi_n],
// which is the same sequence (K K = K(dt) exactly in the phases) with half the transforms.
// The stored array is chi transformed along x (X = DFT_x chi), so a plain step is two sweeps:
//     CK  columns: y forward, K(dt), y inverse                 [load scaled by s]
//     RV  rows:    x inverse, V, sum |.|^2, x forward           (s = 1 / sqrt(sum * dA))
// At the stopping rule's checkpoints the state psi_n itself is formed and measured:
//     CH  columns: y forward, K(dt/2), sum k^2 |psi_k|^2, y inverse
//     RS  rows:    x inverse, sums |psi|^2 and V |psi|^2, x forward
//     CB  columns: y forward, K(dt/2), y inverse
//     RV  rows:    as above
// so E_n (grid_wf_energy_fft, src/grid_wf_fft.c:101-129, plus the potential term) is the energy
// of exactly the state psi_n, and a converged solve ends with psi_n in real space.  The rule
// (host/relax_rule.h) is the one of the host relaxation.
//
// Vector extrapolation, also inside the launch (the rule of grid_wf_gpe_relax_extrapolated):
// every `stride` steps the normalised chi is snapshotted; with the last two differences d1, d2
// parallel (|<d1|d2>| > 0.999 |d1||d2|) and r = Re<d1|d2> / |d1|^2 in (0.05, 0.995), the slowest
// mode's remaining tail d2 r / (1 - r) is added at once, the state renormalised and the rule
// restarted with r as its decay floor.  The snapshots live in per-solve device arrays in the
// x-transformed representation: DFT_x is linear and, up to the factor n, unitary, so
// differences, ratios and the parallelism test are those of real space.

#pragma once
#include "bufio.hpp"
#include "cx.hpp"
#include "grid_hip.h"
#include "rt.hpp"
#define LG_RELAX_FN __device__ static inline
#include "../host/relax_rule.h"

namespace lg {
namespace b2d {

// (threads per pencil, values per thread) for each grid size and tile variant
template <int N, int VAR> struct Tile;
template <> struct Tile<64, 0> { static constexpr int TPP = 8, R = 8; };
template <> struct Tile<96, 0> { static constexpr int TPP = 4, R = 24; };
template <> struct Tile<128, 0> { static constexpr int TPP = 8, R = 16; };
template <> struct Tile<192, 0> { static constexpr int TPP = 8, R = 24; };
template <> struct Tile<256, 0> { static constexpr int TPP = 16, R = 16; };
template <> struct Tile<64, 1> { static constexpr int TPP = 4, R = 16; };
template <> struct Tile<96, 1> { static constexpr int TPP = 4, R = 24; };
template <> struct Tile<128, 1> { static constexpr int TPP = 8, R = 16; };
template <> struct Tile<192, 1> { static constexpr int TPP = 8, R = 24; };
template <> struct Tile<256, 1> { static constexpr int TPP = 8, R = 32; };

template <int N, int VAR> struct Geo {
  static constexpr int TPP = Tile<N, VAR>::TPP, R = Tile<N, VAR>::R;
  static_assert(R * TPP == N && R % TPP == 0, "bad tile");
  static constexpr int PPW = 64 / TPP;  // pencils per wave
  static_assert(N % PPW == 0, "pencil groups must tile the grid");
  static constexpr int GROUPS = N / PPW;            // pencil groups per sweep
  static constexpr int XCH = PPW * R * (TPP + 1);  // floats of one wave's exchange (one real half)
};

template <int VAR> struct MaxXch {
  static constexpr int a = Geo<64, VAR>::XCH, b = Geo<96, VAR>::XCH, c = Geo<128, VAR>::XCH,
                       d = Geo<192, VAR>::XCH, e = Geo<256, VAR>::XCH;
  static constexpr int ab = a > b ? a : b, cd = c > d ? c : d, abcd = ab > cd ? ab : cd;
  static constexpr int value = abcd > e ? abcd : e;
};

constexpr int kMaxN = 256;
constexpr double kExtrapCos = 0.999;

template <typename T> __device__ __forceinline__ T *uniform(T *p) {
  const unsigned long u = (unsigned long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (T *)(((unsigned long)hi << 32) | lo);
}

// Orders the LDS accesses of one wave (its lanes execute together; LDS operations of a wave
// complete in issue order, so only the compiler has to be kept from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One N-point transform of the pencil whose lane t holds x[t + TPP j] in v[j]; xw is the
// pencil's part of its wave's exchange region.  Inverses use conj(DFT(conj x)).
template <int N, int VAR, bool INV, int R = (Tile<N, VAR>::R)>
__device__ __forceinline__ void wfft(float2 (&v)[R], int t, float *xw, const float2 *tw) {
  constexpr int TPP = Geo<N, VAR>::TPP, A = R / TPP;
  if constexpr (INV) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cconj(v[j]);
  }
  dft<R>(v);
  if constexpr (TPP > 1) {
    // W_N^{t k1}: t k1 <= (TPP - 1)(R - 1) < N
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) v[k1] = cmul(v[k1], tw[t * k1]);
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) xw[k1 * (TPP + 1) + t] = v[k1].x;
    wave_sync();
    float ur[A][TPP];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int s = 0; s < TPP; ++s) ur[a][s] = xw[(t + TPP * a) * (TPP + 1) + s];
    wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) xw[k1 * (TPP + 1) + t] = v[k1].y;
    wave_sync();
#pragma unroll
    for (int a = 0; a < A; ++a) {
      float2 u[TPP];
#pragma unroll
      for (int s = 0; s < TPP; ++s) u[s] = make_float2(ur[a][s], xw[(t + TPP * a) * (TPP + 1) + s]);
      dft<TPP>(u);
#pragma unroll
      for (int k2 = 0; k2 < TPP; ++k2) v[a + A * k2] = u[k2];
    }
    wave_sync();  // the wave's next exchange reuses the region
  }
  if constexpr (INV) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cconj(v[j]);
  }
}

// Sum over the 64 lanes of a wave, in a fixed order, as a uniform value: butterfly swaps
// within each half by ds_swizzle (xor masks 1..16 encoded in the instruction, so no lane
// address is kept live), then the two halves from lanes 0 and 32.
__device__ __forceinline__ double wave_sum(double x) {
  static_for<0, 5>([&](auto kk) {
    constexpr int pat = ((1 << decltype(kk)::value) << 10) | 0x1f;  // bit mode: xor mask, and mask 31
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(unsigned)u, pat);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(unsigned)(u >> 32), pat);
    x += __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
  });
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long a = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 0) << 32) |
                               (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, 0);
  const unsigned long long b = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 32) << 32) |
                               (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, 32);
  return __builtin_bit_cast(double, a) + __builtin_bit_cast(double, b);
}

__device__ __forceinline__ double uniform_f64(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Deterministic workgroup sum of K FP64 values, returned as a uniform (scalar) value, so the
// solve's control state derived from it stays in scalar registers.
template <int NT, int K> __device__ __forceinline__ void wg_sum(double (&a)[K], double *red) {
#pragma unroll
  for (int i = 0; i < K; ++i) a[i] = wave_sum(a[i]);
  const int wave = threadIdx.x / 64;
  __syncthreads();
  if (threadIdx.x % 64 == 0) {
#pragma unroll
    for (int i = 0; i < K; ++i) red[K * wave + i] = a[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += red[K * w + i];
    a[i] = uniform_f64(s);
  }
}

// The thread index as a value the compiler cannot see through: each sweep derives its lane
// addresses from it afresh instead of the step loop keeping every sweep's addresses live
// (they cost registers across the whole solve and ended up in scratch).
__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

enum {
  OP_CK = 0,     // columns: y forward, K(dt), y inverse
  OP_CH = 1,     // columns: y forward, K(dt/2) (unless the state is psi_0), sum k^2|.|^2, y inverse
  OP_CB = 2,     // columns: y forward, K(dt/2), y inverse
  OP_RV = 3,     // rows: x inverse, V, sum |.|^2, x forward
  OP_RS = 4,     // rows: x inverse, sums |.|^2 and V|.|^2, x forward
  OP_FIRST = 5,  // rows: sum |.|^2, x forward (the start state)
  OP_LAST = 6    // rows: x inverse, scale (back to real space)
};

// One sweep over every pencil of one axis.  Element j of lane t's pencil: columns ->
// (y = t + TPP j, x = pencil), rows -> (y = pencil, x = t + TPP j).  Loads are scaled by
// `scale` (column sweeps) or stores are (OP_LAST); acc0 / acc1 receive per-thread partial sums.
template <int N, int VAR, int NT, int OP>
__device__ __forceinline__ void sweep(const hgrid_b2d_job &jb, const float2 *tw, const float *kh, const float *kf,
                                      float *xch, float scale, bool use_k, double &acc0, double &acc1) {
  using G = Geo<N, VAR>;
  constexpr int TPP = G::TPP, R = G::R, PPW = G::PPW, NW = NT / 64;
  constexpr bool COLS = (OP == OP_CK || OP == OP_CH || OP == OP_CB);
  constexpr unsigned ESTEP = COLS ? (unsigned)(TPP * N) : (unsigned)TPP;  // elements per j
  const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
  int q, t;
  if constexpr (COLS) {
    q = lane % PPW;  // consecutive lanes: consecutive columns
    t = lane / PPW;
  } else {
    t = lane % TPP;  // consecutive lanes: consecutive x in one row
    q = lane / TPP;
  }
  float *xw = xch + wave * MaxXch<VAR>::value + q * (R * (TPP + 1));
  const auto rpsi = make_rsrc(jb.psi, (unsigned long)N * N * 8);
  const auto rpv = make_rsrc(jb.pv, (unsigned long)N * N * 4);
  const auto rv = make_rsrc(jb.v, (unsigned long)N * N * 8);
  const auto rk2 = make_rsrc(jb.k2, (unsigned long)N * 8);
#pragma unroll 1
  for (int g = wave; g < G::GROUPS; g += NW) {
    const int pen = g * PPW + q;
    const unsigned e0 = COLS ? (unsigned)(t * N + pen) : (unsigned)(pen * N + t);
    float2 v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = bload<float2>(rpsi, e0 * 8u, (unsigned)j * ESTEP * 8u);
    if constexpr (COLS) {
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cscale(v[j], scale);
      wfft<N, VAR, false>(v, t, xw, tw);
      // k-space: kx = pencil, ky = t + TPP j
      const float *kt = (OP == OP_CK) ? kf : kh;
      if (OP != OP_CH || use_k) {
        const float kx = kt[pen];
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = cscale(v[j], kx * kt[t + TPP * j]);
      }
      if constexpr (OP == OP_CH) {
        const double k2x = bload_f64(rk2, (unsigned)pen * 8u, 0);
#pragma unroll
        for (int j = 0; j < R; ++j) acc0 += (k2x + bload_f64(rk2, (unsigned)t * 8u, (unsigned)j * TPP * 8u)) * cnorm2(v[j]);
      }
      wfft<N, VAR, true>(v, t, xw, tw);
    } else if constexpr (OP == OP_FIRST) {
#pragma unroll
      for (int j = 0; j < R; ++j) acc0 += cnorm2(v[j]);
      wfft<N, VAR, false>(v, t, xw, tw);
    } else {
      wfft<N, VAR, true>(v, t, xw, tw);
      if constexpr (OP == OP_RV) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          v[j] = cscale(v[j], bload_f32(rpv, e0 * 4u, (unsigned)j * ESTEP * 4u));
          acc0 += cnorm2(v[j]);
        }
      } else if constexpr (OP == OP_RS) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const double d = cnorm2(v[j]);
          acc0 += d;
          acc1 += d * bload_f64(rv, e0 * 8u, (unsigned)j * ESTEP * 8u);
        }
      } else {  // OP_LAST
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = cscale(v[j], scale);
      }
      if constexpr (OP != OP_LAST) wfft<N, VAR, false>(v, t, xw, tw);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) bstore<float2>(v[j], rpsi, e0 * 8u, (unsigned)j * ESTEP * 8u);
  }
  __syncthreads();  // the next sweep reads what other waves wrote (same workgroup, same CU)
}

// Snapshot of the normalised state c = s X: with have >= 1, d2 = c - s1 and its square norm;
// with have >= 2 also <d1|d2>; d1 <- d2 and s1 <- c.  Elementwise (any mapping).
template <int N, int NT>
__device__ __forceinline__ void snapshot(const hgrid_b2d_job &jb, float s, int have, double &n22, double &xr,
                                         double &xi) {
  const auto rpsi = make_rsrc(jb.psi, (unsigned long)N * N * 8);
  const auto rs1 = make_rsrc(jb.snap, (unsigned long)N * N * 8);
  const auto rd1 = make_rsrc(jb.diff, (unsigned long)N * N * 8);
#pragma unroll 4
  for (int e = opaque_tid(); e < N * N; e += NT) {
    const unsigned off = (unsigned)e * 8u;
    const float2 c = cscale(bload<float2>(rpsi, off, 0), s);
    if (have >= 1) {
      const float2 d2 = csub(c, bload<float2>(rs1, off, 0));
      n22 += cnorm2(d2);
      if (have >= 2) {
        const float2 d1 = bload<float2>(rd1, off, 0);
        xr += (double)d1.x * d2.x + (double)d1.y * d2.y;
        xi += (double)d1.x * d2.y - (double)d1.y * d2.x;
      }
      bstore<float2>(d2, rd1, off, 0);
    }
    bstore<float2>(c, rs1, off, 0);
  }
  __syncthreads();
}

// X <- s X + c d (d: the newest difference), with its square norm.
template <int N, int NT>
__device__ __forceinline__ void extrapolate(const hgrid_b2d_job &jb, float s, float c, double &nrm) {
  const auto rpsi = make_rsrc(jb.psi, (unsigned long)N * N * 8);
  const auto rd1 = make_rsrc(jb.diff, (unsigned long)N * N * 8);
#pragma unroll 4
  for (int e = opaque_tid(); e < N * N; e += NT) {
    const unsigned off = (unsigned)e * 8u;
    const float2 x = cscale(bload<float2>(rpsi, off, 0), s), d = bload<float2>(rd1, off, 0);
    const float2 y = make_float2(fmaf(c, d.x, x.x), fmaf(c, d.y, x.y));
    nrm += cnorm2(y);
    bstore<float2>(y, rpsi, off, 0);
  }
  __syncthreads();
}

// The stopping rule and the extrapolation's rule restart run out of line: their FP64 exp / log
// constants would otherwise be hoisted out of the step loop and held in registers throughout.
__device__ __noinline__ int rule_feed(lg_relax_state *rs, long n, double e, double tol) {
  return lg_relax_feed(rs, n, e, tol);
}
__device__ __noinline__ void rule_restart(lg_relax_state *rs, double r, int stride) {
  lg_relax_init(rs);
  rs->lnq_floor = log(r) / (double)stride;
}

template <int N, int VAR, int NT>
__device__ void solve(const hgrid_b2d_job &jb, float *xch, float2 *tw, float *kh, float *kf,
                                   double *red) {
  const unsigned long long t_start = wall_clock64();
  for (int m = threadIdx.x; m < N; m += NT) {
    tw[m] = ((const float2 *)jb.tw)[m];
    kh[m] = jb.kh[m];
    kf[m] = jb.kf[m];
  }
  __syncthreads();
  const double dA = jb.dA;
  double e = 0.0, x0 = 0.0;
  float s;
  {
    double a[1] = {0.0};
    sweep<N, VAR, NT, OP_FIRST>(jb, tw, kh, kf, xch, 1.f, false, a[0], x0);
    wg_sum<NT, 1>(a, red);
    s = (float)(1.0 / sqrt(a[0] * dA));
  }
  lg_relax_state rs;
  lg_relax_init(&rs);
  int n = 0, n0 = 0, have = 0, nextrap = 0, n_done = jb.max_steps;
  double n11 = 0.0;
  bool is_chi = false;  // false while the stored state is psi_0 itself (no kinetic factor pending)
#pragma unroll 1
  for (;; ++n) {
    const bool last = n >= jb.max_steps;
    if (jb.stride > 0 && n > 0 && !last && n % jb.stride == 0) {
      double a[3] = {0.0, 0.0, 0.0};
      snapshot<N, NT>(jb, s, have, a[0], a[1], a[2]);
      wg_sum<NT, 3>(a, red);
      if (have < 2) {
        n11 = a[0];
        ++have;
      } else {
        const double n22 = a[0], r = a[1] / n11;
        if (!isfinite(n22) || !isfinite(a[1]) || !isfinite(a[2])) {
          n_done = -2;  // the host reports the solve as diverged
          break;
        }
        if (n11 > 0.0 && n22 > 0.0 && sqrt(a[1] * a[1] + a[2] * a[2]) > kExtrapCos * sqrt(n11 * n22) && r > 0.05 &&
            r < 0.995) {
          double b[1] = {0.0};
          extrapolate<N, NT>(jb, s, (float)(r / (1.0 - r)), b[0]);
          wg_sum<NT, 1>(b, red);
          // sum |X|^2 = N sum |chi|^2 along x
          s = (float)(1.0 / sqrt(b[0] / N * dA));
          have = 0;
          ++nextrap;
          rule_restart(&rs, r, jb.stride);
          n0 = n;
        } else {
          n11 = n22;
        }
      }
    }
    const bool check = last || (long)(n - n0) == rs.next;
    if (check) {
      double a[3] = {0.0, 0.0, 0.0};
      sweep<N, VAR, NT, OP_CH>(jb, tw, kh, kf, xch, s, is_chi, a[0], x0);
      sweep<N, VAR, NT, OP_RS>(jb, tw, kh, kf, xch, 1.f, false, a[1], a[2]);
      wg_sum<NT, 3>(a, red);
      e = (jb.tfac * a[0] + a[2]) / a[1];
      if (n < jb.hist_len && threadIdx.x == 0) jb.hist[n] = e;
      const int verdict = last ? 0 : rule_feed(&rs, (long)(n - n0), e, jb.tol);
      if (verdict != 0 || last) {
        if (verdict > 0) n_done = n;
        if (verdict < 0) n_done = -2;
        double d0 = 0.0, d1 = 0.0;
        sweep<N, VAR, NT, OP_LAST>(jb, tw, kh, kf, xch, (float)(1.0 / ((double)N * sqrt(a[1] * dA))), false, d0, d1);
        break;
      }
      sweep<N, VAR, NT, OP_CB>(jb, tw, kh, kf, xch, (float)(1.0 / sqrt(a[1] * dA)), true, x0, x0);
    } else {
      if (n < jb.hist_len && threadIdx.x == 0) jb.hist[n] = __builtin_nan("");
      sweep<N, VAR, NT, OP_CK>(jb, tw, kh, kf, xch, s, true, x0, x0);
    }
    double a[1] = {0.0};
    sweep<N, VAR, NT, OP_RV>(jb, tw, kh, kf, xch, 1.f, false, a[0], x0);
    wg_sum<NT, 1>(a, red);
    s = (float)(1.0 / sqrt(a[0] * dA));
    is_chi = true;
  }
  const unsigned long long t_end = wall_clock64();
  if (threadIdx.x == 0) {
    jb.out[0] = e;
    jb.out[1] = (double)n;
    jb.out[2] = (double)n_done;
    jb.out[3] = (double)t_start;
    jb.out[4] = (double)t_end;
    jb.out[5] = (double)nextrap;
  }
}

// MW workgroups per CU: (NT / 64) MW / 4 waves per SIMD, which sets the register budget.
template <int VAR, int NT, int MW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu((NT / 64) * MW / 4)))
batch2d_kernel(const hgrid_b2d_job *jobs, int njobs, int *queue) {
  __shared__ float xch[(NT / 64) * MaxXch<VAR>::value];
  __shared__ float2 s_tw[kMaxN];
  __shared__ float s_kh[kMaxN], s_kf[kMaxN];
  __shared__ double red[3 * (NT / 64)];
  __shared__ int next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(queue, 1);
    __syncthreads();
    // uniform across the workgroup: the job descriptor is then read with scalar loads
    const int j = __builtin_amdgcn_readfirstlane(next);
    if (j >= njobs) break;  // every wave of the workgroup leaves together
    hgrid_b2d_job jb = jobs[j];
    jb.psi = uniform(jb.psi);
    jb.pv = uniform(jb.pv);
    jb.v = uniform(jb.v);
    jb.kh = uniform(jb.kh);
    jb.kf = uniform(jb.kf);
    jb.k2 = uniform(jb.k2);
    jb.tw = uniform(jb.tw);
    jb.snap = uniform(jb.snap);
    jb.diff = uniform(jb.diff);
    jb.out = uniform(jb.out);
    jb.hist = uniform(jb.hist);
    jb.hist_len = __builtin_amdgcn_readfirstlane(jb.hist_len);
    jb.n = __builtin_amdgcn_readfirstlane(jb.n);
    jb.max_steps = __builtin_amdgcn_readfirstlane(jb.max_steps);
    jb.stride = __builtin_amdgcn_readfirstlane(jb.stride);
    switch (jb.n) {
      case 64: solve<64, VAR, NT>(jb, xch, s_tw, s_kh, s_kf, red); break;
      case 96: solve<96, VAR, NT>(jb, xch, s_tw, s_kh, s_kf, red); break;
      case 128: solve<128, VAR, NT>(jb, xch, s_tw, s_kh, s_kf, red); break;
      case 192: solve<192, VAR, NT>(jb, xch, s_tw, s_kh, s_kf, red); break;
      case 256: solve<256, VAR, NT>(jb, xch, s_tw, s_kh, s_kf, red); break;
      default:
        if (threadIdx.x == 0) {
          jb.out[0] = 0.0;
          jb.out[1] = -1.0;
          jb.out[2] = -1.0;
          jb.out[3] = 0.0;
          jb.out[4] = 0.0;
          jb.out[5] = 0.0;
        }
    }
  }
}

template <int VAR, int NT, int MW> int launch(const hgrid_b2d_job *jobs, int njobs, int *queue, int nblocks) {
  hipLaunchKernelGGL((batch2d_kernel<VAR, NT, MW>), dim3(nblocks), dim3(NT), 0, stream(), jobs, njobs, queue);
  return check_launch("batch2d");
}

// per-variant entry points (one translation unit each)
int launch_v0(const hgrid_b2d_job *jobs, int njobs, int *queue, int nthreads, int per_cu, int nblocks);
int launch_v1(const hgrid_b2d_job *jobs, int njobs, int *queue, int nthreads, int per_cu, int nblocks);

}  // namespace b2d
}  // namespace lg
