// Batched 2-D imaginary-time ground states in one persistent launch (complex64 storage,
// FP64 reductions).
//
// Each workgroup takes whole solves from an atomic work queue (the host orders the queue by
// expected cost, largest first) and relaxes one n x n grid at a time with workgroup barriers
// only: a grid of at most 256^2 complex64 values (512 KiB) stays resident in L2 / Infinity
// Cache while its workgroup sweeps it.  One imaginary-time step of
//     psi <- IFFT[K FFT[ P_V IFFT[K FFT[psi]] ]],  K = exp(-dt/2 k^2 / 2m) / n (per axis),
//     P_V = exp(-dt V)
// is four sweeps, because the data stays transformed along x between steps:
//     C1  columns: forward y, K (and sum k^2|psi_k|^2 for T), inverse y    [load scaled by s]
//     R2  rows:    inverse x, P_V, forward x
//     C2  columns: forward y, K, inverse y
//     R3  rows:    inverse x, sums |psi|^2 and V|psi|^2 (next state), forward x
// The state is renormalised every step through the load scale s = 1/sqrt(norm * dA) of C1,
// E_n = T_n + s_n^2 dA sum V|psi|^2 is complete after C1 of step n, and the solve stops when
// its energy is converged to tol by the windowed rule of host/relax_rule.h (or at
// max_steps).  A final row pass returns the normalised state to real space.  The work per pencil reuses the register/LDS pencil transform of fft_pass.hpp;
// the tile shape (threads per pencil, pencils per sweep, workgroup size and workgroups per
// CU) is a template/launch choice made by the autotuner (host/batch2d.c).
// Reference operators composed: grid_wf_propagate (src/grid_wf.c:379-432) in imaginary time,
// grid_wf_normalize (:694-700), grid_wf_energy_fft (src/grid_wf_fft.c:101-129).

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
template <> struct Tile<64, 1> { static constexpr int TPP = 4, R = 16; };
template <> struct Tile<96, 0> { static constexpr int TPP = 4, R = 24; };
template <> struct Tile<96, 1> { static constexpr int TPP = 4, R = 24; };
template <> struct Tile<128, 0> { static constexpr int TPP = 8, R = 16; };
template <> struct Tile<128, 1> { static constexpr int TPP = 4, R = 32; };
template <> struct Tile<192, 0> { static constexpr int TPP = 8, R = 24; };
template <> struct Tile<192, 1> { static constexpr int TPP = 8, R = 24; };
template <> struct Tile<256, 0> { static constexpr int TPP = 16, R = 16; };
template <> struct Tile<256, 1> { static constexpr int TPP = 8, R = 32; };

template <int N, int VAR, int NT> struct Geo {
  static constexpr int TPP = Tile<N, VAR>::TPP, R = Tile<N, VAR>::R;
  static_assert(R * TPP == N && R % TPP == 0, "bad tile");
  static constexpr int P = NT / TPP;             // pencil slots per sweep chunk (all threads busy)
  static constexpr int CHUNKS = (N + P - 1) / P;  // slots past the last pencil idle
  static constexpr int LDS = P * R * (TPP + 1);   // floats, one real half of the exchange
};

template <int VAR, int NT> struct MaxLds {
  static constexpr int a = Geo<64, VAR, NT>::LDS, b = Geo<96, VAR, NT>::LDS, c = Geo<128, VAR, NT>::LDS,
                       d = Geo<192, VAR, NT>::LDS, e = Geo<256, VAR, NT>::LDS;
  static constexpr int ab = a > b ? a : b, cd = c > d ? c : d, abcd = ab > cd ? ab : cd;
  static constexpr int value = abcd > e ? abcd : e;
};

template <typename T> __device__ __forceinline__ T *uniform(T *p) {
  const unsigned long u = (unsigned long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (T *)(((unsigned long)hi << 32) | lo);
}

enum { OP_C1 = 0, OP_R2 = 1, OP_C2 = 2, OP_R3 = 3, OP_FIRST = 4, OP_LAST = 5 };

// One N-point transform of the pencil owned by (q, t): thread t holds x[t + TPP j].
template <int N, int VAR, int NT, int R = (Tile<N, VAR>::R)>
__device__ __forceinline__ void pencil_fft(float2 (&v)[R], int t, int q, bool inverse, const double2 *tw, float *lds) {
  using G = Geo<N, VAR, NT>;
  constexpr int TPP = G::TPP;
  if (inverse) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cconj(v[j]);
  }
  dft<R>(v);
  if constexpr (TPP > 1) {
    // W_N^{t k1} by an FP64 recurrence from the exact table entry W_N^t (error < R * 1e-16,
    // far below FP32 rounding), instead of R - 1 table loads with live addresses
    const double2 w1 = tw[t];
    double2 w = w1;
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) {
      v[k1] = cmul(v[k1], make_float2((float)w.x, (float)w.y));
      w = cmul(w, w1);
    }
    constexpr int A = R / TPP;
    const int base = q * (R * (TPP + 1));
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) lds[base + k1 * (TPP + 1) + t] = v[k1].x;
    __syncthreads();
    float ur[A][TPP];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int s = 0; s < TPP; ++s) ur[a][s] = lds[base + (t + TPP * a) * (TPP + 1) + s];
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) lds[base + k1 * (TPP + 1) + t] = v[k1].y;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < A; ++a) {
      float2 u[TPP];
#pragma unroll
      for (int s = 0; s < TPP; ++s) u[s] = make_float2(ur[a][s], lds[base + (t + TPP * a) * (TPP + 1) + s]);
      dft<TPP>(u);
#pragma unroll
      for (int k2 = 0; k2 < TPP; ++k2) v[a + A * k2] = u[k2];
    }
  }
  if (inverse) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cconj(v[j]);
  }
}

// Deterministic workgroup sum of two FP64 values (result broadcast to every thread).
template <int NT> __device__ __forceinline__ void wg_sum2(double &a, double &b, double *red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int wave = threadIdx.x / 64;
  __syncthreads();
  if (threadIdx.x % 64 == 0) {
    red[2 * wave] = a;
    red[2 * wave + 1] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    sa += red[2 * w];
    sb += red[2 * w + 1];
  }
  a = sa;
  b = sb;
}

// One sweep over all pencils of the grid.  Returns per-thread partial sums in acc0/acc1.
// Element j of a thread's pencil: columns -> (y = t + TPP j, x = pen), rows -> (y = pen,
// x = t + TPP j); the per-element byte step is a compile-time constant (SGPR offset).
template <int N, int VAR, int NT, int OP>
__device__ __forceinline__ void sweep(const hgrid_b2d_job &jb, float scale, double &acc0, double &acc1, float *lds) {
  using G = Geo<N, VAR, NT>;
  constexpr int TPP = G::TPP, R = G::R, P = G::P;
  constexpr bool COLS = (OP == OP_C1 || OP == OP_C2);
  constexpr unsigned ESTEP = COLS ? (unsigned)(TPP * N) : (unsigned)TPP;  // elements per j
  const int tid = threadIdx.x;
  int q, t;
  if constexpr (COLS) {
    q = tid % P;  // consecutive lanes: consecutive columns (coalesced rows of P values)
    t = tid / P;
  } else {
    t = tid % TPP;  // consecutive lanes: consecutive x in one row
    q = tid / TPP;
  }
  const float *kt = jb.kt;
  const double *k2 = jb.k2;
  const double2 *tw = (const double2 *)jb.tw;
  const auto rpsi = make_rsrc(jb.psi, (unsigned long)N * N * 8);
  const auto rpv = make_rsrc(jb.pv, (unsigned long)N * N * 4);
  const auto rv = make_rsrc(jb.v, (unsigned long)N * N * 8);
  const auto rkt = make_rsrc(kt, (unsigned long)N * 4);
  const auto rk2 = make_rsrc(k2, (unsigned long)N * 8);
#pragma unroll 1
  for (int c = 0; c < G::CHUNKS; ++c) {
    const int pen = c * P + q;
    const bool active = pen < N;
    const unsigned e0 = COLS ? (unsigned)(t * N + pen) : (unsigned)(pen * N + t);
    float2 v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = active ? bload<float2>(rpsi, e0 * 8u, (unsigned)j * ESTEP * 8u) : make_float2(0.f, 0.f);
    if constexpr (OP == OP_C1 || OP == OP_C2) {
      if (OP == OP_C1) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = cscale(v[j], scale);
      }
      pencil_fft<N, VAR, NT>(v, t, q, false, tw, lds);
      __builtin_amdgcn_sched_barrier(0);
      // (kx = pen, ky = t + TPP j)
      const float kx = bload_f32(rkt, (unsigned)pen * 4u, 0);
      if (OP == OP_C1) {
        const double k2x = bload_f64(rk2, (unsigned)pen * 8u, 0);
#pragma unroll
        for (int j = 0; j < R; ++j) acc0 += (k2x + bload_f64(rk2, (unsigned)t * 8u, (unsigned)j * TPP * 8u)) * cnorm2(v[j]);
      }
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cscale(v[j], kx * bload_f32(rkt, (unsigned)t * 4u, (unsigned)j * TPP * 4u));
      pencil_fft<N, VAR, NT>(v, t, q, true, tw, lds);
    } else if constexpr (OP == OP_R2) {
      pencil_fft<N, VAR, NT>(v, t, q, true, tw, lds);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cscale(v[j], bload_f32(rpv, e0 * 4u, (unsigned)j * ESTEP * 4u));
      pencil_fft<N, VAR, NT>(v, t, q, false, tw, lds);
    } else if constexpr (OP == OP_R3 || OP == OP_FIRST) {
      if (OP == OP_R3) pencil_fft<N, VAR, NT>(v, t, q, true, tw, lds);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const double d = cnorm2(v[j]);
        acc0 += d;
        acc1 += d * bload_f64(rv, e0 * 8u, (unsigned)j * ESTEP * 8u);
      }
      pencil_fft<N, VAR, NT>(v, t, q, false, tw, lds);
    } else {  // OP_LAST: back to real space, normalised
      pencil_fft<N, VAR, NT>(v, t, q, true, tw, lds);
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cscale(v[j], scale);
    }
    if (active) {
#pragma unroll
      for (int j = 0; j < R; ++j) bstore<float2>(v[j], rpsi, e0 * 8u, (unsigned)j * ESTEP * 8u);
    }
  }
  __syncthreads();  // the next sweep reads what this one wrote (same workgroup, same CU cache)
}

template <int N, int VAR, int NT> __device__ void solve(const hgrid_b2d_job &jb, float *lds, double *red) {
  const unsigned long long t_start = wall_clock64();
  double a0 = 0.0, a1 = 0.0;
  // x-transform the start state; its norm and potential sum give s_0 and V_0
  sweep<N, VAR, NT, OP_FIRST>(jb, 1.f, a0, a1, lds);
  wg_sum2<NT>(a0, a1, red);
  double norm = a0, vsum = a1;
  double e = 0.0;
  lg_relax_state rs;
  lg_relax_init(&rs);
  int n_done = jb.max_steps;
  int step = 0;
#pragma unroll 1
  for (; step < jb.max_steps; ++step) {
    const double s = 1.0 / sqrt(norm * jb.dA);
    double t0 = 0.0, t1 = 0.0;
    sweep<N, VAR, NT, OP_C1>(jb, (float)s, t0, t1, lds);
    wg_sum2<NT>(t0, t1, red);
    // E of the normalised state that entered this step
    e = jb.tfac * t0 + s * s * jb.dA * vsum;
    if (step < jb.hist_len && threadIdx.x == 0) jb.hist[step] = e;
    const int verdict = lg_relax_feed(&rs, step, e, jb.tol);
    if (verdict < 0) {
      n_done = -2;  // non-finite energy: the host reports the solve as diverged
      break;
    }
    if (verdict > 0) {
      n_done = step;
      // finish this step so the state is whole, then stop
      double d0 = 0.0, d1 = 0.0;
      sweep<N, VAR, NT, OP_R2>(jb, 1.f, d0, d1, lds);
      sweep<N, VAR, NT, OP_C2>(jb, 1.f, d0, d1, lds);
      a0 = a1 = 0.0;
      sweep<N, VAR, NT, OP_R3>(jb, 1.f, a0, a1, lds);
      wg_sum2<NT>(a0, a1, red);
      norm = a0;
      ++step;
      break;
    }
    double d0 = 0.0, d1 = 0.0;
    sweep<N, VAR, NT, OP_R2>(jb, 1.f, d0, d1, lds);
    sweep<N, VAR, NT, OP_C2>(jb, 1.f, d0, d1, lds);
    a0 = a1 = 0.0;
    sweep<N, VAR, NT, OP_R3>(jb, 1.f, a0, a1, lds);
    wg_sum2<NT>(a0, a1, red);
    norm = a0;
    vsum = a1;
  }
  {
    double d0 = 0.0, d1 = 0.0;
    // the stored data is x-transformed: the inverse brings a factor N, the norm a factor s
    sweep<N, VAR, NT, OP_LAST>(jb, (float)(1.0 / (sqrt(norm * jb.dA) * N)), d0, d1, lds);
  }
  const unsigned long long t_end = wall_clock64();
  if (threadIdx.x == 0) {
    jb.out[0] = e;
    jb.out[1] = (double)step;
    jb.out[2] = (double)n_done;
    jb.out[3] = (double)t_start;
    jb.out[4] = (double)t_end;
  }
}

template <int VAR, int NT, int MW>
__global__ void __launch_bounds__(NT, MW) batch2d_kernel(const hgrid_b2d_job *jobs, int njobs, int *queue) {
  __shared__ float lds[MaxLds<VAR, NT>::value];
  __shared__ double red[2 * (NT / 64)];
  __shared__ int next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(queue, 1);
    __syncthreads();
    // uniform across the workgroup: the job descriptor is then read with scalar loads
    const int j = __builtin_amdgcn_readfirstlane(next);
    if (j >= njobs) break;  // every wave of the workgroup leaves together
    hgrid_b2d_job jb = jobs[j];
    // the descriptor is uniform: keep its fields in scalar registers
    jb.psi = uniform(jb.psi);
    jb.pv = uniform(jb.pv);
    jb.v = uniform(jb.v);
    jb.kt = uniform(jb.kt);
    jb.k2 = uniform(jb.k2);
    jb.tw = uniform(jb.tw);
    jb.out = uniform(jb.out);
    jb.hist = uniform(jb.hist);
    jb.hist_len = __builtin_amdgcn_readfirstlane(jb.hist_len);
    jb.n = __builtin_amdgcn_readfirstlane(jb.n);
    jb.max_steps = __builtin_amdgcn_readfirstlane(jb.max_steps);
    switch (jb.n) {
      case 64: solve<64, VAR, NT>(jb, lds, red); break;
      case 96: solve<96, VAR, NT>(jb, lds, red); break;
      case 128: solve<128, VAR, NT>(jb, lds, red); break;
      case 192: solve<192, VAR, NT>(jb, lds, red); break;
      case 256: solve<256, VAR, NT>(jb, lds, red); break;
      default:
        if (threadIdx.x == 0) {
          jb.out[0] = 0.0;
          jb.out[1] = -1.0;
          jb.out[2] = -1.0;
          jb.out[3] = 0.0;
          jb.out[4] = 0.0;
        }
    }
  }
}

template <int VAR, int NT, int MW> int launch(const hgrid_b2d_job *jobs, int njobs, int *queue, int nblocks) {
  hipLaunchKernelGGL((batch2d_kernel<VAR, NT, MW>), dim3(nblocks), dim3(NT), 0, stream(), jobs, njobs, queue);
  return check_launch("batch2d");
}

// per-variant entry points (one translation unit each)
int launch_v0(const hgrid_b2d_job *jobs, int njobs, int *queue, int nthreads, int nblocks);
int launch_v1(const hgrid_b2d_job *jobs, int njobs, int *queue, int nthreads, int nblocks);

}  // namespace b2d
}  // namespace lg
