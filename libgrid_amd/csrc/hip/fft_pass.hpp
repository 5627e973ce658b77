// Pencil FFT passes for gfx950: one pass transforms every pencil of one axis of a 3-D
// complex array and may fuse pointwise work in front of, between and behind its transforms.
//
// Decomposition of one pencil of length L = R * TPP (TPP threads per pencil, R values per
// thread).  Thread t holds x[t + TPP*j], j < R.  With n = t + TPP*j and k = k1 + R*k2:
//     X[k1 + R*k2] = sum_t W_TPP^{t*k2} * ( W_L^{t*k1} * sum_j x[t + TPP*j] W_R^{j*k1} )
// stage 1 is an R-point DFT in registers, then a twiddle W_L^{t*k1} from an FP64-accurate
// table, an LDS transpose, and stage 2: R/TPP TPP-point DFTs per thread.  The result lands
// in the same distribution as the input (thread t holds X[t + TPP*m]), so a second
// transform (inverse after forward, or forward after inverse) can follow in registers.
//
// Memory layout (squeezed shape n0 x n1 x n2, n2 contiguous):
//   axis 2 ("contiguous" pass): pencil r = i0*n1 + i1 starts at r*L, stride 1.  Lanes of one
//            pencil are adjacent, so one wave reads TPP*sizeof(C) runs of consecutive bytes.
//   axis 1/0 ("strided" pass): pencils are flattened over (outer, i2); a workgroup takes P
//            consecutive i2 values so a wave load covers P*sizeof(C) contiguous bytes per row.
// Global traffic uses raw buffer instructions: one 32-bit VGPR offset per thread plus a
// uniform SGPR offset per element, instead of a 64-bit address per element.
//
// Pass kinds (compile time, so each kernel carries only the code it runs):
//   KIND_FFT : one transform; optional pending scale on load, |psi|^2 / reductions on the
//              input, real post factor and |psi|^2 reduction on the output
//   KIND_KIN : forward, complex kinetic factor (optionally sum k^2|psi_k|^2 first), inverse
//   KIND_KIR : KIND_KIN with real factors (imaginary time).  Separate kinds because a kernel
//              holding both factor paths pays the registers of both (strided 512-point FP32
//              kinetic pass: 254 VGPRs with both, 166 with the complex path alone)
//   KIND_KIB : both factor paths, chosen at run time, for the FP32 real-factor passes the
//              compiler builds worse as KIND_KIR (kib_length): the strided 512-point one came
//              out at 256 VGPRs + 50 spilled to AGPRs (1 wave per SIMD, 1.00 ms at 512^3) as
//              KIND_KIR and at 254 VGPRs (2 waves, 0.72 ms) here
//   KIND_POT : inverse, mid * exp(c (cons + V)) with complex V, forward
//   KIND_GPE : inverse, mid * exp(c (cons + vext + g rho)), forward
//   KIND_GPR : KIND_GPE with real c (imaginary time) only: without the complex-phase code the
//              256-point FP64 pass fits 3 waves per SIMD (166 instead of 224 VGPRs)
//   KIND_PRE : one transform on a contiguous axis, then |psi|^2 / V_eff / reductions of the
//              result, nothing stored (the start-of-step pass over a state kept transformed)
//   KIND_FY  : KIND_FFT with at most the pending scale, the post factor and the output norm
//              (the y passes of a step, plain transforms): without the input-side paths the
//              kernel keeps fewer registers (strided 512-point FP32: 210 -> 112 VGPRs)
// Inverse transforms use conj(DFT(conj(x))), so one set of butterflies serves both signs.
// Reference operators this replaces: cgrid_fft / cgrid_inverse_fft (src/cgrid.c:1822-1846,
// src/cgrid_fft.c:101-146), the kinetic phase (src/grid_wf_fft.c:141-193), the potential
// phase (src/grid_wf.c:569-600), the 1/N scaling (src/cgrid.c:1857-1876) and the reductions
// (src/cgrid.c:1610-1629, src/cgrid_diff.c:817-879, src/ngrid.c:355-376).

#pragma once
#include "cx.hpp"
#include "grid_hip.h"
#include "bufio.hpp"
#include "phase_math.hpp"
#include "rt.hpp"

namespace lg {

enum { KIND_FFT = 0, KIND_KIN = 1, KIND_POT = 2, KIND_GPE = 3, KIND_PRE = 4, KIND_GPR = 5, KIND_KIR = 6, KIND_KIB = 7,
       KIND_FY = 8 };


// Threads per pencil (TPP) and values per thread (R) for each supported length L = R * TPP:
// powers of two up to 1024 and 3 * 2^k up to 384.  Two register stages (R-point DFTs, then
// R / TPP TPP-point DFTs per thread) need R to be a multiple of TPP.  Contiguous 512-point
// pencils instead run three 8-point stages with 64 threads per pencil (THREE): a pencil is then
// one wave, which orders its exchanges itself, and keeps 8 values per thread instead of 32 (the
// two-stage 16 x 32 split ran the contiguous GPE pass at 1 wave/SIMD).  The same form for
// contiguous 256-point pencils (8 x 8 x 4, 32 threads) was measured too: it cut the FP64 passes'
// registers (GPE 234 -> 152) but made the 256^3 GPE step slower (0.822 against 0.816 ms: the GPE
// pass 124 -> 136 us), so 256 keeps the 16 x 16 split.  Strided 512-point pencils keep the
// two-stage split as well: their three-stage form spans 16 waves per workgroup, and its eight
// workgroup barriers per two-transform pass cost more than the registers save (x kinetic pass
// at 512^3 FP32: 1156 us three-stage, 810 us two-stage).
template <int L, bool CONTIG> struct Split {
  static constexpr bool THREE = L == 512 && CONTIG;
  static constexpr int TPP = THREE ? 64
                             : (L <= 16 || L == 24) ? 1
                             : L == 32 || L == 48 || L == 96 ? 4
                             : L == 64 || L == 128 || L == 192 || L == 384 ? 8
                             : L == 256 || L == 512 ? 16
                             : 32;
  static constexpr int R = L / TPP;
  static_assert(R * TPP == L, "unsupported length");
  static_assert(THREE || R % TPP == 0, "stage 2 needs R to be a multiple of TPP");
  static_assert(!THREE || (R == 8 && TPP == 64), "the three-stage split is 8 x 8 x 8");
};

#define LG_FOR_EACH_LENGTH(X) \
  X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(1024)

// Wide tiles: on the outermost axis of a 512-point FP64 grid a row of a tile is 8 pencils x 16 B
// = 128 B, one DRAM burst per 4 MiB-strided row.  There the workgroup takes 16 pencils (256 B
// per row): x kinetic pass at 512^3 1.29 -> 0.97 ms.  The y axis (8 KiB row stride) keeps 8
// pencils and two workgroups per CU: with 16 its pass went 0.77 -> 0.84 ms.
template <typename C, int L, bool CONTIG> constexpr bool wide_capable() { return !CONTIG && L == 512 && sizeof(C) == 16; }

template <typename C, int L, bool CONTIG, bool WIDE = false> struct Geom {
  static constexpr int TPP = Split<L, CONTIG>::TPP;
  static constexpr int R = Split<L, CONTIG>::R;
  static constexpr bool THREE = Split<L, CONTIG>::THREE;
  // workgroup size: two-stage lengths take 256 threads unless one real half of the LDS exchange
  // exceeds 32 KiB (and the tiles are not wide).  Three-stage (contiguous only): pencils are one
  // wave each, four per workgroup.
  static constexpr int NT = THREE || WIDE ? 256 : (TPP > 1 && (long)R * (long)sizeof(C) * 128 > 32768) ? 128 : 256;
  static constexpr int P = NT / TPP;  // pencils per workgroup
  // complex entries per pencil in the three-stage exchange buffer (512 + 4: consecutive pencils
  // start 4 entries apart in the LDS banks)
  static constexpr int PS3 = 516;
  // two-stage: the exchange moves real parts, then imaginary parts, through one real buffer;
  // three-stage: interleaved complex entries (LDS sizes are counted in reals)
  static constexpr int LDS_CONTIG = THREE ? 2 * P * PS3 : TPP > 1 ? P * R * (TPP + 1) : 1;
  static constexpr int LDS_STRIDED = THREE ? 2 * P * PS3 : TPP > 1 ? P * R * (TPP + 1) : 1;
};

// Positions in a pencil's exchange buffer of the three-stage transform (complex entries).
// Exchange 1 moves Y[t1 + 8 t2][k1]: written by thread t1 + 8 t2 (k1 = 0..7), read by thread
// t1 + 8 k1 (t2 = 0..7); exchange 2 moves Z[t1, k1][k2]: written by thread t1 + 8 k1 (k2 = 0..7),
// read by thread k1 + 8 k2 (t1 = 0..7).  The XOR terms permute entries within rows of 8 so that
// the lanes an LDS cycle serves (8 to 32 of them for 8- or 16-byte entries at 128 or 256 bytes
// per cycle) hit distinct banks in all four patterns; an earlier permutation left 2-way
// conflicts in exchange 2's reads (33.6 M conflict cycles per 512^3 FP64 GPE pass).
__device__ __forceinline__ int x1_index(int k1, int t2, int t1) { return (k1 << 6) + ((t2 ^ (k1 & 3)) << 3) + t1; }
__device__ __forceinline__ int x2_index(int k1, int k2, int t1) {
  return (k2 << 6) + (k1 << 3) + (t1 ^ ((k1 ^ (k2 & 3)) & 7));
}

// Ordering of a pencil's LDS exchange.  A contiguous three-stage pencil is exactly one wave,
// whose LDS accesses execute in order: only the compiler must be kept from moving them across
// this point.  Otherwise the pencil spans the waves of the workgroup.
template <bool WAVE> __device__ __forceinline__ void pencil_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// The twiddle table W_L^m, m < L, as a buffer resource whose address is hidden from the
// optimiser: the twiddles of a second transform in the same pass are then reloaded (L1 hits)
// instead of being kept live across the mid-pass work, and, as buffer loads, they are counted
// by vmcnt alone, so waiting for them does not also wait for the LDS exchange (flat loads of a
// laundered pointer count on both)
template <typename C, int L> __device__ __forceinline__ __amdgpu_buffer_rsrc_t twiddle_rsrc(const C *tw) {
  unsigned long long b = (unsigned long long)tw;
  __asm__ volatile("" : "+s"(b));
  return make_rsrc((const void *)b, (unsigned long)L * sizeof(C));
}

// One 512-point transform of the pencil owned by (q, t), 64 threads of 8 values, thread t holding
// x[t + 64 j].  With n = t1 + 8 t2 + 64 j and k = k1 + 8 k2 + 64 k3:
//   stage 1: 8-point DFTs over j, twiddle W_512^{t k1}        (thread t = t1 + 8 t2)
//   stage 2: 8-point DFTs over t2, twiddle W_64^{t1 k2}       (thread t1 + 8 k1)
//   stage 3: 8-point DFTs over t1                             (thread k1 + 8 k2)
// so thread t ends holding X[t + 64 k3]: the input distribution.  Both twiddles come from the
// length-512 table (W_64^m = W_512^{8m}).
template <typename C, bool CONTIG>
__device__ __forceinline__ void fft_core3(C (&v)[8], int t, int q, bool inverse, const C *__restrict__ tw,
                                          typename real_of<C>::type *ldsr) {
  constexpr int PS = Geom<C, 512, CONTIG>::PS3;
  C *lds = reinterpret_cast<C *>(ldsr) + q * PS;
  if (inverse) {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = cconj(v[j]);
  }
  const auto twr = twiddle_rsrc<C, 512>(tw);
  dft<8>(v);
#pragma unroll
  for (int k1 = 1; k1 < 8; ++k1) v[k1] = cmul(v[k1], bload<C>(twr, (unsigned)(t * k1) * (unsigned)sizeof(C), 0));
  pencil_sync<CONTIG>();  // the previous transform's exchange 2 may still be reading
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) lds[x1_index(k1, t >> 3, t & 7)] = v[k1];
  pencil_sync<CONTIG>();
  const int a = t & 7, b = t >> 3;  // stage 2: (t1, k1) = (a, b); stage 3: (k1, k2) = (a, b)
#pragma unroll
  for (int t2 = 0; t2 < 8; ++t2) v[t2] = lds[x1_index(b, t2, a)];
  dft<8>(v);
#pragma unroll
  for (int k2 = 1; k2 < 8; ++k2) v[k2] = cmul(v[k2], bload<C>(twr, (unsigned)(8 * a * k2) * (unsigned)sizeof(C), 0));
  pencil_sync<CONTIG>();  // exchange 1's reads of this buffer are done
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) lds[x2_index(b, k2, a)] = v[k2];
  pencil_sync<CONTIG>();
#pragma unroll
  for (int t1 = 0; t1 < 8; ++t1) v[t1] = lds[x2_index(a, b, t1)];
  dft<8>(v);
  if (inverse) {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = cconj(v[j]);
  }
}

template <bool CONTIG, int R, int TPP, int P>
__device__ __forceinline__ int lds_index(int q, int k1, int t) {
  if constexpr (CONTIG) return q * (R * (TPP + 1)) + k1 * (TPP + 1) + t;
  // strided: a row of TPP + 1 slots of P pencils per k1.  The spare slot staggers the rows, so
  // that the reads of stage 2 (one k1 row per thread t) from the lanes of different t fall in
  // different banks (at 512 points the rows without it cost 2-4 conflict cycles per access)
  else return (k1 * (TPP + 1) + t) * P + q;
}

// One L-point transform in two register stages (the decomposition at the top of this file).
template <typename C, int L, bool CONTIG, bool WIDE>
__device__ __forceinline__ void fft_core2(C (&v)[(Split<L, CONTIG>::R)], int t, int q, bool inverse, const C *__restrict__ tw,
                                          typename real_of<C>::type *lds) {
  constexpr int TPP = Split<L, CONTIG>::TPP, R = Split<L, CONTIG>::R, P = Geom<C, L, CONTIG, WIDE>::P;
  static_assert(!CONTIG || 64 % TPP == 0, "a contiguous pencil's threads lie in one wave");
  if (inverse) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cconj(v[j]);
  }
  dft<R>(v);
  if constexpr (TPP > 1) {
    const auto twr = twiddle_rsrc<C, L>(tw);
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) v[k1] = cmul(v[k1], bload<C>(twr, (unsigned)(t * k1) * (unsigned)sizeof(C), 0));
    constexpr int A = R / TPP;
    // real parts.  A contiguous pencil's TPP threads are consecutive lanes of one wave and its
    // exchange region is its own, so ordering that wave's LDS accesses is enough there.
    pencil_sync<CONTIG>();  // the previous exchange may still be reading
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) lds[lds_index<CONTIG, R, TPP, P>(q, k1, t)] = v[k1].x;
    pencil_sync<CONTIG>();
    typename real_of<C>::type ur[A][TPP];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int s = 0; s < TPP; ++s) ur[a][s] = lds[lds_index<CONTIG, R, TPP, P>(q, t + TPP * a, s)];
    pencil_sync<CONTIG>();
    // imaginary parts
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) lds[lds_index<CONTIG, R, TPP, P>(q, k1, t)] = v[k1].y;
    pencil_sync<CONTIG>();
#pragma unroll
    for (int a = 0; a < A; ++a) {
      C u[TPP];
#pragma unroll
      for (int s = 0; s < TPP; ++s) u[s] = mk<C>(ur[a][s], lds[lds_index<CONTIG, R, TPP, P>(q, t + TPP * a, s)]);
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

// |k|^2 of element i of an L-point axis with k spacing dk: k = i dk (i <= L/2) or (i - L) dk,
// squared with its own rounding (no contraction into the sum that uses it), as the host tables
// of |k|^2 are formed
template <int L> __device__ __forceinline__ double k2_of(int i, double dk) {
#pragma clang fp contract(off)
  const double k = (double)(i <= L / 2 ? i : i - L) * dk;
  return k * k;
}

// ---- one L-point transform of the pencil owned by (q, t) ----------------------------------

template <typename C, int L, bool CONTIG, bool WIDE>
__device__ __forceinline__ void fft_core(C (&v)[(Split<L, CONTIG>::R)], int t, int q, bool inverse, const C *__restrict__ tw,
                                         typename real_of<C>::type *lds) {
  if constexpr (Split<L, CONTIG>::THREE) {
    fft_core3<C, CONTIG>(v, t, q, inverse, tw, lds);
  } else {
    fft_core2<C, L, CONTIG, WIDE>(v, t, q, inverse, tw, lds);
  }
}


// exp(c * V) with complex c and complex V (FP64 argument).
template <typename C> __device__ __forceinline__ C phase_factor(double cr, double ci, double vr, double vi);

template <> __device__ __forceinline__ double2 phase_factor<double2>(double cr, double ci, double vr, double vi) {
  const double ar = cr * vr - ci * vi, ai = cr * vi + ci * vr;
  double s, c;
  sincos_f64(ai, &s, &c);
  const double m = exp_f64(ar);
  return make_double2(m * c, m * s);
}

// exp(x) for a real exponent, in FP64 for either storage precision: an FP32-rounded factor
// would be the same per element every step, i.e. a fixed perturbation of V by about
// 6e-8 / dt, which moves the FP32 fixed point of imaginary-time relaxation by ~1e-9 Eh.
// Applied with cscale_d, only the product is rounded, and that rounding follows the data.
__device__ __forceinline__ double real_decay(double x) { return exp_f64(x); }

template <> __device__ __forceinline__ float2 phase_factor<float2>(double cr, double ci, double vr, double vi) {
  // FP64 argument, reduced by multiples of pi/2 in FP64, then FP32 sin / cos / exp
  const double ar = cr * vr - ci * vi, ai = cr * vi + ci * vr;
  float s, c;
  sincos_f32(ai, &s, &c);
  const float m = expf((float)ar);
  return make_float2(m * c, m * s);
}

// exp(i x) (FP64 argument): the factor exp(c V) of real time (c_re = 0) with a real V, whose
// exponent c V has no real part.  Equal to phase_factor(0, c_im, V, 0) bit for bit, without
// evaluating exp(0).
template <typename C> __device__ __forceinline__ C unit_phase(double x);
template <> __device__ __forceinline__ double2 unit_phase<double2>(double x) {
  double s, c;
  sincos_f64(x, &s, &c);
  return make_double2(c, s);
}
template <> __device__ __forceinline__ float2 unit_phase<float2>(double x) {
  float s, c;
  sincos_f32(x, &s, &c);
  return make_float2(c, s);
}

// C is the arithmetic type, S the storage type of psi (and of a complex potential): S = C, or
// S = float2 with C = double2 ("FP32 storage, FP64 arithmetic": only the stores round).
// Waves per SIMD asked of the register allocator (minimum, maximum), per kernel, where measured
// faster; (1, 8) leaves the allocator its own choice.  Strided kinetic passes in FP64: at 256
// points exactly 3 (158 VGPRs: 100 -> 92 us at 256^3; at 4 waves the allocator's 128-VGPR
// schedule ran 111 us), at 512 points at least 2 (256 VGPRs instead of 256 + AGPR spills and 1
// wave: 1.43 -> 1.29 ms at 512^3).  Hints that forced scratch spills (start-of-step pass at 4,
// GPE pass at 3) were slower and are not used.
template <typename C, int L, bool CONTIG, int KIND> constexpr int waves_hint() {
  if ((KIND == KIND_KIN || KIND == KIND_KIR) && !CONTIG && L == 512 && sizeof(C) == 16) return 2;
  return ((KIND == KIND_KIN || KIND == KIND_KIR) && !CONTIG && L == 256 && sizeof(C) == 16) ? 3 : 1;
}
template <typename C, int L, bool CONTIG, int KIND> constexpr int waves_cap() {
  return ((KIND == KIND_KIN || KIND == KIND_KIR) && !CONTIG && L == 256 && sizeof(C) == 16) ? 3 : 8;
}

template <typename C, int L, bool CONTIG, int KIND, typename S = C, bool WIDE = false>
__global__ void __launch_bounds__((Geom<C, L, CONTIG, WIDE>::NT))
__attribute__((amdgpu_waves_per_eu(waves_hint<C, L, CONTIG, KIND>(), waves_cap<C, L, CONTIG, KIND>()))) fft_pass_kernel(const hgrid_pass a) {
  using Rl = typename real_of<C>::type;
  using G = Geom<C, L, CONTIG, WIDE>;
  constexpr int TPP = Split<L, CONTIG>::TPP, R = Split<L, CONTIG>::R, NT = G::NT, P = G::P;
  constexpr int LDS_N = CONTIG ? G::LDS_CONTIG : G::LDS_STRIDED;
  constexpr unsigned CS = sizeof(S);  // bytes per stored element
  __shared__ __align__(16) Rl lds[LDS_N];
  __shared__ double red_lds[NT / 64][HGRID_RED_NCHAN];
  // kinetic pass on a strided axis: the axis's factor table staged in LDS, published by the
  // transform's first workgroup barrier, or by one of its own where the transform has none
  // (one thread per pencil)
  constexpr bool KT_LDS = (KIND == KIND_KIN || KIND == KIND_KIR || KIND == KIND_KIB) && !CONTIG;
  __shared__ __align__(16) double2 kt_lds[KT_LDS ? L : 1];

  const int tid = threadIdx.x;
  int q, t;
  if constexpr (CONTIG) {
    t = tid % TPP;
    q = tid / TPP;
  } else {
    q = tid % P;
    t = tid / P;
  }
  const long n0 = a.n0, n1 = a.n1, n2 = a.n2;
  const long ntot = n0 * n1 * n2;
  const long npencils = ntot / L;
  // XCD-aware order on the y axis and on the wide x tiles: workgroup b runs on XCD b % 8, so
  // the first 8 (nb / 8) workgroups are renumbered to give each XCD a contiguous run of tiles
  // (the tail keeps its number).  At 512^3 FP64 the y pass went from 916 to 780 us and the wide
  // x kinetic pass from 984-988 to 953-971 us (two boxes, alternating runs).  The other x passes
  // (256^3 FP64 kinetic: 93.0-93.3 -> 95.8-96.3 us; 512^3 FP32: 648 -> 654 us) and the
  // contiguous axis were slower in this order, so they keep the hardware order.
  long blk = blockIdx.x;
  if constexpr (!CONTIG) {
    if (a.axis == 1 || WIDE) {
      const long nb = gridDim.x, per = nb / 8;
      if (blk < 8 * per) blk = (blk % 8) * per + blk / 8;
    }
  }
  const long r = blk * P + q;
  const bool active = r < npencils;
  long base, c1, c2;
  unsigned astride;
  if constexpr (CONTIG) {
    base = r * L;
    astride = 1;
    c1 = r / n1;
    c2 = r - c1 * n1;
  } else {
    if (a.axis == 1) {
      const long o = r / n2, z = r - o * n2;
      base = o * n1 * n2 + z;
      astride = (unsigned)n2;
      c1 = o;
      c2 = z;
    } else {
      base = r;
      astride = (unsigned)(n1 * n2);
      c1 = r / n2;
      c2 = r - c1 * n2;
    }
  }
  if (!active) {
    base = 0;
    c1 = c2 = 0;
  }
  // element j of this thread: flat index base + (t + TPP*j)*astride
  const unsigned e0 = (unsigned)base + (unsigned)t * astride;  // flat index of element 0
  const unsigned estep = (unsigned)TPP * astride;              // flat index step per j
  // KIND_PRE: a KIND_FFT pass whose other flags are known to be clear, compiled without them
  constexpr unsigned PRE_KEEP = HGRID_PF_PRESCALE | HGRID_PF_PRE_RHO | HGRID_PF_PRE_RED | HGRID_PF_PRE_VEFF;
  const unsigned fl = KIND == KIND_PRE  ? (a.flags & PRE_KEEP) | HGRID_PF_PRE_AFTER | HGRID_PF_NOSTORE
                      : KIND == KIND_FY ? a.flags & (HGRID_PF_PRESCALE | HGRID_PF_POST_RED)
                                        : a.flags;
  const C *__restrict__ tw = (const C *)a.tw;
  const auto rin = make_rsrc(a.in, (unsigned long)ntot * CS);
  if constexpr (KT_LDS) {
    const double2 *ka = (const double2 *)a.ktab_a;
    for (int i = tid; i < L; i += NT) kt_lds[i] = ka[i];
  }

  // Loads by every thread, zeroed by a select for a thread past the last pencil (which reads
  // elements of pencil 0, base 0, in bounds), instead of a branch around each load.  Measured
  // against the branch form: contiguous passes faster (256^3 FP64 start-of-step pass 101.7-102.0
  // -> 100.3-100.5 us, potential pass 114.7-115.0 -> 113.1-113.2 us), 512-point plain transforms
  // too (on y: FP64 779 -> 766 us, FP32 399 -> 394 us), but the wide FP64 x kinetic pass slower
  // (966-970 -> 995-1004 us) and the 256-point y pass a little slower (81.9-82.0 -> 82.4-82.6
  // us), so those keep the branches.
  constexpr bool SELECT_LOADS = CONTIG || (KIND == KIND_FY && L == 512 && !WIDE);
  C v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if constexpr (SELECT_LOADS) {
      const C x = widen<C>(bload<S>(rin, e0 * CS, (unsigned)j * estep * CS));
      v[j] = active ? x : mk<C>(0, 0);
    } else {
      v[j] = active ? widen<C>(bload<S>(rin, e0 * CS, (unsigned)j * estep * CS)) : mk<C>(0, 0);
    }
  }
  if constexpr (KT_LDS && TPP == 1) __syncthreads();

  double acc[HGRID_RED_NCHAN] = {0.0, 0.0, 0.0, 0.0, 0.0};

  if constexpr (KIND == KIND_FFT || KIND == KIND_POT || KIND == KIND_PRE || KIND == KIND_FY) {
    if (fl & HGRID_PF_PRESCALE) {
      const C s = mk<C>((Rl)a.scale[0], (Rl)a.scale[1]);
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cmul(v[j], s);
    }
  }
  // density / V_eff / input reductions of the real-space data: before the transform, or after
  // it with PRE_AFTER (the data is then stored transformed along this axis and the pass
  // inverse-transforms it in registers); element j keeps flat index e0 + j * estep either way
  const auto rvx = make_rsrc(a.vext, a.vext ? (unsigned long)ntot * 8 : 0);
  // the start-of-step pass loads vext with the state, before its transform, so the two reads
  // share one memory latency (256^3 FP64: 112 -> 102 us at 164 instead of 128 VGPRs)
  double vxp[KIND == KIND_PRE ? R : 1];
  if constexpr (KIND == KIND_PRE) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      // a null vext has a zero-length buffer, whose loads return 0
      const double x = bload_f64(rvx, (e0 + (unsigned)j * estep) * 8u, 0);
      vxp[j] = active ? x : 0.0;
    }
  }
  auto pre_ops = [&]() {
    const auto rrho = make_rsrc(a.rho_out, (unsigned long)ntot * 8);
    if constexpr (KIND == KIND_PRE) {
      // vext is held: the stores go under one branch, the sums need none (an inactive thread
      // holds zeros)
      if (active && (fl & (HGRID_PF_PRE_VEFF | HGRID_PF_PRE_RHO))) {
        const bool veff = fl & HGRID_PF_PRE_VEFF;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const double d = cnorm2(v[j]);
          bstore_f64(veff ? vxp[j] + a.g * d : d, rrho, (e0 + (unsigned)j * estep) * 8u, 0);
        }
      }
      if (fl & HGRID_PF_PRE_RED) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const double d = cnorm2(v[j]);
          acc[HGRID_RED_NORM] += d;
          acc[HGRID_RED_QUART] += d * d;
          acc[HGRID_RED_VEXT] += vxp[j] * d;
        }
      }
      return;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const unsigned e = e0 + (unsigned)j * estep;
      const double d = cnorm2(v[j]);
      const double vx = active ? bload_f64(rvx, e * 8u, 0) : 0.0;
      if (active) {
        if (fl & HGRID_PF_PRE_VEFF) bstore_f64(vx + a.g * d, rrho, e * 8u, 0);
        else if (fl & HGRID_PF_PRE_RHO) bstore_f64(d, rrho, e * 8u, 0);
      }
      if (fl & HGRID_PF_PRE_RED) {
        acc[HGRID_RED_NORM] += d;
        acc[HGRID_RED_QUART] += d * d;
        acc[HGRID_RED_VEXT] += vx * d;
      }
    }
  };
  constexpr unsigned PRE_ANY = HGRID_PF_PRE_RHO | HGRID_PF_PRE_RED | HGRID_PF_PRE_VEFF;
  if constexpr (KIND == KIND_FFT) {
    if ((fl & PRE_ANY) && !(fl & HGRID_PF_PRE_AFTER)) pre_ops();
  }

  if constexpr (KIND == KIND_POT) {
    // the 4th-order step's outer V stages: potential factor before or after one transform
    if (!(fl & HGRID_PF_POT_NOFFT1)) fft_core<C, L, CONTIG, WIDE>(v, t, q, a.dir1 > 0, tw, lds);
  } else {
    fft_core<C, L, CONTIG, WIDE>(v, t, q, a.dir1 > 0, tw, lds);
  }

  if constexpr (KIND == KIND_FFT || KIND == KIND_PRE) {
    if ((fl & PRE_ANY) && (fl & HGRID_PF_PRE_AFTER)) pre_ops();
  }

  if constexpr (KIND == KIND_FFT || KIND == KIND_KIN || KIND == KIND_KIR || KIND == KIND_KIB) {
    if (fl & HGRID_PF_MID_KRED) {
      // |k|^2 of the pass axis formed in registers, bit for bit the table entry: a table load
      // per element, issued among the 32 live values of a 512-point pencil, came out serialised
      // (one wait per load; 0.75 -> 0.90 ms for the FP32 512^3 kinetic pass)
      const double kk = a.k2_1[c1] + a.k2_2[c2];
#pragma unroll
      for (int j = 0; j < R; ++j) acc[HGRID_RED_KIN] += (kk + k2_of<L>(t + TPP * j, a.k2_dk)) * cnorm2(v[j]);
    }
  }
  if constexpr (KIND == KIND_KIN || KIND == KIND_KIR || KIND == KIND_KIB) {
    // FP64 tables whatever the storage precision; the factor is formed in FP64.  Real
    // factors (imaginary time) multiply in FP64 and only the product is rounded, so FP32
    // storage sees no fixed per-element error in K; complex ones are rounded once to C.
    const double2 *ka = KT_LDS ? (const double2 *)kt_lds : (const double2 *)a.ktab_a;
    const double2 f = cmul(((const double2 *)a.ktab_1)[c1], ((const double2 *)a.ktab_2)[c2]);
    auto real_factors = [&]() {
      if constexpr (std::is_same<C, float2>::value) {
        // FP32 data: groups of four behind scheduling fences keep few FP64 factors live
#pragma unroll
        for (int j0 = 0; j0 < R; j0 += 4) {
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int j = j0; j < (j0 + 4 < R ? j0 + 4 : R); ++j) v[j] = cscale_d(v[j], ka[t + TPP * j].x * f.x);
        }
        __builtin_amdgcn_sched_barrier(0);
      } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = cscale_d(v[j], ka[t + TPP * j].x * f.x);
      }
    };
    auto complex_factors = [&]() {
      // complex factors (real time): formed in FP64 in groups of four behind scheduling fences,
      // so that few FP64 table entries and products are live at once
#pragma unroll
      for (int j0 = 0; j0 < R; j0 += 4) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = j0; j < (j0 + 4 < R ? j0 + 4 : R); ++j) v[j] = cmul(v[j], round_to<C>(cmul(ka[t + TPP * j], f)));
      }
      __builtin_amdgcn_sched_barrier(0);
    };
    if constexpr (KIND == KIND_KIR) {
      real_factors();
    } else if constexpr (KIND == KIND_KIN) {
      complex_factors();
    } else {
      if (fl & HGRID_PF_KIN_REAL) real_factors();
      else complex_factors();
    }
    if (fl & HGRID_PF_KCUT) {
      // CFFT: components with |k - k0|^2 >= kmax^2 become exact zeros
      const double kc1 = a.kc_1[c1], kc2 = a.kc_2[c2];
#pragma unroll
      for (int j = 0; j < R; ++j)
        if ((a.kc_a[t + TPP * j] + kc1) + kc2 >= a.kcut2) v[j] = mk<C>(0, 0);
    }
  } else if constexpr (KIND == KIND_POT || KIND == KIND_GPE || KIND == KIND_GPR) {
    const Rl mid = (Rl)a.mid;
    const auto rpot = make_rsrc(a.pot, (unsigned long)ntot * CS);
    const auto rvx = make_rsrc(a.vext, a.vext ? (unsigned long)ntot * 8 : 0);
    const auto rrho = make_rsrc(a.rho, a.rho ? (unsigned long)ntot * 8 : 0);
    if constexpr (KIND == KIND_POT) {
      // loads first, then the factors in groups of four behind scheduling fences, so that few
      // exp / sincos evaluations are live at once (234 -> 178 VGPRs at L = 256 in FP64)
      S pv[R];
#pragma unroll
      for (int j = 0; j < R; ++j) pv[j] = bload<S>(rpot, (e0 + (unsigned)j * estep) * CS, 0);
#pragma unroll
      for (int j0 = 0; j0 < R; j0 += 4) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = j0; j < (j0 + 4 < R ? j0 + 4 : R); ++j)
          v[j] = cmul(cscale(v[j], mid), phase_factor<C>(a.c_re, a.c_im, a.cons + (double)pv[j].x, (double)pv[j].y));
      }
      __builtin_amdgcn_sched_barrier(0);
    } else {
      // real potential: V = vext + g rho, rho alone when it already holds V_eff, or vext alone
      double vr[R];
      if (a.vext && a.rho) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const unsigned e = e0 + (unsigned)j * estep;
          vr[j] = a.cons + bload_f64(rvx, e * 8u, 0) + a.g * bload_f64(rrho, e * 8u, 0);
        }
      } else if (a.vext) {
#pragma unroll
        for (int j = 0; j < R; ++j) vr[j] = a.cons + bload_f64(rvx, (e0 + (unsigned)j * estep) * 8u, 0);
      } else {
#pragma unroll
        for (int j = 0; j < R; ++j) vr[j] = a.cons + bload_f64(rrho, (e0 + (unsigned)j * estep) * 8u, 0);
      }
      if (KIND == KIND_GPR || a.c_im == 0.0) {
        // imaginary time: a real decay factor exp(c_re * V)
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = cscale_d(v[j], (double)mid * real_decay(a.c_re * vr[j]));
      } else if constexpr (KIND != KIND_GPR) {
        if (a.c_re == 0.0) {
#pragma unroll
          for (int j = 0; j < R; ++j) v[j] = cmul(cscale(v[j], mid), unit_phase<C>(a.c_im * vr[j]));
        } else {
#pragma unroll
          for (int j = 0; j < R; ++j) v[j] = cmul(cscale(v[j], mid), phase_factor<C>(a.c_re, a.c_im, vr[j], 0.0));
        }
      }
    }
  }

  if constexpr (KIND == KIND_POT) {
    if (!(fl & HGRID_PF_POT_NOFFT2)) fft_core<C, L, CONTIG, WIDE>(v, t, q, a.dir2 > 0, tw, lds);
  } else if constexpr (KIND != KIND_FFT && KIND != KIND_PRE && KIND != KIND_FY) {
    fft_core<C, L, CONTIG, WIDE>(v, t, q, a.dir2 > 0, tw, lds);
  }

  if constexpr (KIND == KIND_FFT || KIND == KIND_FY) {
    if (a.post != 1.0) {
      const Rl post = (Rl)a.post;
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cscale(v[j], post);
    }
  }
  if constexpr (KIND == KIND_FFT || KIND == KIND_GPE || KIND == KIND_GPR || KIND == KIND_FY) {
    // the stored values (a merged relaxation step renormalises from the output of its potential
    // pass; kinetic and complex-potential passes never reduce their output, and the accumulation
    // would cost the strided 512-point kinetic pass its second wave per SIMD)
    if (fl & HGRID_PF_POST_RED) {
#pragma unroll
      for (int j = 0; j < R; ++j) acc[HGRID_RED_POST] += cnorm2(v[j]);
    }
  }
  if (!(fl & HGRID_PF_NOSTORE) && active) {
    const auto rout = make_rsrc(a.out, (unsigned long)ntot * CS);
#pragma unroll
    for (int j = 0; j < R; ++j) bstore<S>(round_to<S>(v[j]), rout, e0 * CS, (unsigned)j * estep * CS);
  }

  if (fl & (HGRID_PF_PRE_RED | HGRID_PF_MID_KRED | HGRID_PF_POST_RED)) {
    // deterministic workgroup reduction: wave butterflies, then waves in order
#pragma unroll
    for (int c = 0; c < HGRID_RED_NCHAN; ++c) {
      double x = acc[c];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
      acc[c] = x;
    }
    const int wave = tid / 64, lane = tid % 64;
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < HGRID_RED_NCHAN; ++c) red_lds[wave][c] = acc[c];
    }
    __syncthreads();
    if (tid < HGRID_RED_NCHAN) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) s += red_lds[w][tid];
      a.partials[(long)tid * gridDim.x + blockIdx.x] = s;
    }
  }
}

template <typename C, int L, bool CONTIG, int KIND, typename S, bool WIDE> int launch_geom(const hgrid_pass *p) {
  constexpr int NT = Geom<C, L, CONTIG, WIDE>::NT, P = Geom<C, L, CONTIG, WIDE>::P;
  const long npencils = (p->n0 * p->n1 * p->n2) / L;
  const long nblocks = (npencils + P - 1) / P;
  if (nblocks <= 0) return 0;
  hipLaunchKernelGGL((fft_pass_kernel<C, L, CONTIG, KIND, S, WIDE>), dim3((unsigned)nblocks), dim3(NT), 0, stream(), *p);
  return check_launch("fft_pass");
}

template <typename C, int L, bool CONTIG, int KIND, typename S> int launch(const hgrid_pass *p) {
  if constexpr (wide_capable<C, L, CONTIG>()) {
    if (p->axis == 0) return launch_geom<C, L, CONTIG, KIND, S, true>(p);
  }
  return launch_geom<C, L, CONTIG, KIND, S, false>(p);
}

template <typename C, bool CONTIG, int KIND, typename S> int launch_len(const hgrid_pass *p, long L) {
#define LG_CASE(n) \
  case n: return launch<C, n, CONTIG, KIND, S>(p);
  switch (L) {
    LG_FOR_EACH_LENGTH(LG_CASE)
    default: return -1;
  }
#undef LG_CASE
}

// the lengths whose FP32 real-factor kinetic pass runs with both factor paths (KIND_KIB): there
// the real-only kernel comes out at fewer waves per SIMD (compiler resource report: 6, 8, 16
// and 24 points one wave fewer, 512 and 1024 one wave instead of two)
constexpr bool kib_length(long L) { return L == 6 || L == 8 || L == 16 || L == 24 || L == 512 || L == 1024; }

template <typename C, bool CONTIG, typename S> int launch_kib(const hgrid_pass *p, long L) {
  switch (L) {
    case 6: return launch<C, 6, CONTIG, KIND_KIB, S>(p);
    case 8: return launch<C, 8, CONTIG, KIND_KIB, S>(p);
    case 16: return launch<C, 16, CONTIG, KIND_KIB, S>(p);
    case 24: return launch<C, 24, CONTIG, KIND_KIB, S>(p);
    case 512: return launch<C, 512, CONTIG, KIND_KIB, S>(p);
    case 1024: return launch<C, 1024, CONTIG, KIND_KIB, S>(p);
    default: return -1;
  }
}

template <typename C, typename S = C> int launch_prec(const hgrid_pass *p, long L, int kind) {
  const bool contig = p->axis == 2;
  switch (kind) {
    case KIND_FFT: return contig ? launch_len<C, true, KIND_FFT, S>(p, L) : launch_len<C, false, KIND_FFT, S>(p, L);
    case KIND_FY: return contig ? launch_len<C, true, KIND_FY, S>(p, L) : launch_len<C, false, KIND_FY, S>(p, L);
    case KIND_KIN: return contig ? launch_len<C, true, KIND_KIN, S>(p, L) : launch_len<C, false, KIND_KIN, S>(p, L);
    case KIND_KIR: return contig ? launch_len<C, true, KIND_KIR, S>(p, L) : launch_len<C, false, KIND_KIR, S>(p, L);
    case KIND_KIB:
      if constexpr (std::is_same<C, float2>::value) {
        return contig ? launch_kib<C, true, S>(p, L) : launch_kib<C, false, S>(p, L);
      }
      return -1;
    case KIND_POT: return contig ? launch_len<C, true, KIND_POT, S>(p, L) : -1;
    case KIND_GPE: return contig ? launch_len<C, true, KIND_GPE, S>(p, L) : -1;
    case KIND_GPR: return contig ? launch_len<C, true, KIND_GPR, S>(p, L) : -1;
    case KIND_PRE: return contig ? launch_len<C, true, KIND_PRE, S>(p, L) : -1;
    default: return -1;
  }
}

// pencils per workgroup of a pass along `axis` (the geometry launch() picks)
template <typename C, bool CONTIG> long pencils_per_block(long L, int axis) {
#define LG_CASE(n) \
  case n: return wide_capable<C, n, CONTIG>() && axis == 0 ? Geom<C, n, CONTIG, true>::P : Geom<C, n, CONTIG>::P;
  switch (L) {
    LG_FOR_EACH_LENGTH(LG_CASE)
    default: return 0;
  }
#undef LG_CASE
}

// per-precision entry points (one translation unit each, compiled in parallel)
int launch_f64(const hgrid_pass *p, long L, int kind);
int launch_f32(const hgrid_pass *p, long L, int kind);
int launch_f32s_f64(const hgrid_pass *p, long L, int kind);  // float2 storage, FP64 arithmetic

}  // namespace lg
