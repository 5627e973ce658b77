// Device-side complex arithmetic and in-register DFT kernels for gfx950.
//
// The data types are HIP's double2 / float2 (interleaved re/im, the same layout as C99
// `double _Complex` / `float _Complex`).  dft<N>() transforms an array that lives in VGPRs;
// every index is a compile-time constant (static_for), so after inlining the "array" is
// just a set of registers and the butterflies are straight-line FMA code.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace lg {

template <typename C> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };

template <typename C> __device__ __forceinline__ C mk(typename real_of<C>::type x, typename real_of<C>::type y) {
  C r; r.x = x; r.y = y; return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cneg(C a) { return mk<C>(-a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cscale(C a, typename real_of<C>::type s) { return mk<C>(a.x * s, a.y * s); }
// multiply by -i and +i
template <typename C> __device__ __forceinline__ C mul_mi(C a) { return mk<C>(a.y, -a.x); }
template <typename C> __device__ __forceinline__ C mul_pi(C a) { return mk<C>(-a.y, a.x); }
template <typename C> __device__ __forceinline__ double cnorm2(C a) {
  // |a|^2 accumulated in FP64 whatever the storage precision (float*float is exact in double)
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// compile-time loop
template <int I, int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(std::forward<F>(f));
  }
}

// cos(2*pi*m/64), m = 0..16, correctly rounded (computed to 40+ digits).
constexpr double kQuarterCos64[17] = {
    1.0,
    0.99518472667219688624483695310947992157,
    0.98078528040323044912618223613423903697,
    0.95694033573220886493579788698026996948,
    0.92387953251128675612818318939678828682,
    0.88192126434835502971275686366038834950,
    0.83146961230254523707878837761790575673,
    0.77301045336273696081090660975846980097,
    0.70710678118654752440084436210484903928,
    0.63439328416364549821517161322549337067,
    0.55557023301960222474283081394853287437,
    0.47139673682599764855638762590525437765,
    0.38268343236508977172845998403039886676,
    0.29028467725446236763619237581739527469,
    0.19509032201612826784828486847702224092,
    0.09801714032956060199419556388864184586,
    0.0};

__host__ __device__ constexpr double cos64(int m) {
  m &= 63;
  return m <= 16 ? kQuarterCos64[m] : m <= 32 ? -kQuarterCos64[32 - m] : m <= 48 ? -kQuarterCos64[m - 32] : kQuarterCos64[64 - m];
}
__host__ __device__ constexpr double sin64(int m) { return cos64(m - 16); }

// a * W_N^E with W_N = exp(-2 pi i / N) (forward sign), N a power of two <= 64.
template <int N, int E, typename C> __device__ __forceinline__ C twmul(C a) {
  using R = typename real_of<C>::type;
  static_assert(64 % N == 0, "in-register twiddles support N | 64");
  constexpr int m = ((E % N) * (64 / N)) & 63;
  if constexpr (m == 0) return a;
  else if constexpr (m == 16) return mul_mi(a);
  else if constexpr (m == 32) return cneg(a);
  else if constexpr (m == 48) return mul_pi(a);
  else if constexpr (m == 8) {
    constexpr R r = (R)kQuarterCos64[8];
    return mk<C>((a.x + a.y) * r, (a.y - a.x) * r);
  } else if constexpr (m == 24) {
    constexpr R r = (R)kQuarterCos64[8];
    return mk<C>((a.y - a.x) * r, -(a.x + a.y) * r);
  } else if constexpr (m == 40) {
    constexpr R r = (R)kQuarterCos64[8];
    return mk<C>(-(a.x + a.y) * r, (a.x - a.y) * r);
  } else if constexpr (m == 56) {
    constexpr R r = (R)kQuarterCos64[8];
    return mk<C>((a.x - a.y) * r, (a.x + a.y) * r);
  } else {
    constexpr R c = (R)cos64(m);
    constexpr R s = (R)(-sin64(m));
    return mk<C>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// Forward DFT of N points held in registers: X[k] = sum_n x[n] exp(-2 pi i n k / N).
// Natural order in and out.  Radix-4 decimation in time, radix-2 for a leftover factor 2.
template <int N, typename C> __device__ __forceinline__ void dft(C (&x)[N]) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if constexpr (N % 4 == 0) {
    constexpr int M = N / 4;
    C s0[M], s1[M], s2[M], s3[M];
    static_for<0, M>([&](auto mm) {
      constexpr int m = decltype(mm)::value;
      s0[m] = x[4 * m + 0];
      s1[m] = x[4 * m + 1];
      s2[m] = x[4 * m + 2];
      s3[m] = x[4 * m + 3];
    });
    dft<M>(s0);
    dft<M>(s1);
    dft<M>(s2);
    dft<M>(s3);
    static_for<0, M>([&](auto kk) {
      constexpr int k = decltype(kk)::value;
      C b0 = s0[k];
      C b1 = twmul<N, 1 * k>(s1[k]);
      C b2 = twmul<N, 2 * k>(s2[k]);
      C b3 = twmul<N, 3 * k>(s3[k]);
      C t0 = cadd(b0, b2), t1 = csub(b0, b2), t2 = cadd(b1, b3), t3 = mul_mi(csub(b1, b3));
      x[k] = cadd(t0, t2);
      x[k + M] = cadd(t1, t3);
      x[k + 2 * M] = csub(t0, t2);
      x[k + 3 * M] = csub(t1, t3);
    });
  } else {
    static_assert(N % 2 == 0, "power-of-two sizes only in dft<N>");
    constexpr int M = N / 2;
    C s0[M], s1[M];
    static_for<0, M>([&](auto mm) {
      constexpr int m = decltype(mm)::value;
      s0[m] = x[2 * m];
      s1[m] = x[2 * m + 1];
    });
    dft<M>(s0);
    dft<M>(s1);
    static_for<0, M>([&](auto kk) {
      constexpr int k = decltype(kk)::value;
      C b = twmul<N, k>(s1[k]);
      x[k] = cadd(s0[k], b);
      x[k + M] = csub(s0[k], b);
    });
  }
}

}  // namespace lg
