// Device-side complex arithmetic and in-register DFT kernels for gfx950.
//
// The data types are HIP's double2 / float2 (interleaved re/im, the same layout as C99
// `double _Complex` / `float _Complex`).  dft<N>() transforms an array that lives in VGPRs;
// every index is a compile-time constant (static_for), so after inlining the "array" is
// just a set of registers and the butterflies are straight-line FMA code.  N may be any
// divisor of 192 (powers of two up to 64, and 3, 6, 12, 24, 48, 96, 192): radix-4 steps,
// then radix-3, then radix-2.  The functions are also host-callable so that a CPU harness
// can check them against a brute-force DFT.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

#define LG_CX __host__ __device__ __forceinline__

namespace lg {

template <typename C> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };

template <typename C> LG_CX C mk(typename real_of<C>::type x, typename real_of<C>::type y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename C> LG_CX C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> LG_CX C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> LG_CX C cneg(C a) { return mk<C>(-a.x, -a.y); }
template <typename C> LG_CX C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C> LG_CX C cmul(C a, C b) { return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> LG_CX C cscale(C a, typename real_of<C>::type s) { return mk<C>(a.x * s, a.y * s); }
// a * s with the product formed in FP64 and rounded once to a's precision
template <typename C> LG_CX C cscale_d(C a, double s) {
  using R = typename real_of<C>::type;
  return mk<C>((R)((double)a.x * s), (R)((double)a.y * s));
}
// a complex rounded once to the precision of C (identity when the types agree)
template <typename C, typename A> LG_CX C round_to(A k) {
  using R = typename real_of<C>::type;
  return mk<C>((R)k.x, (R)k.y);
}
// a stored value widened to the arithmetic type C (exact)
template <typename C, typename A> LG_CX C widen(A k) {
  using R = typename real_of<C>::type;
  return mk<C>((R)k.x, (R)k.y);
}
// a * k with the product formed in FP64 and rounded once to a's precision (FP64 phase tables
// applied to FP32 data: the table moduli are not rounded step after step)
template <typename C> LG_CX C cmul_d(C a, double2 k) {
  using R = typename real_of<C>::type;
  const double x = (double)a.x, y = (double)a.y;
  return mk<C>((R)(x * k.x - y * k.y), (R)(x * k.y + y * k.x));
}
// multiply by -i and +i
template <typename C> LG_CX C mul_mi(C a) { return mk<C>(a.y, -a.x); }
template <typename C> LG_CX C mul_pi(C a) { return mk<C>(-a.y, a.x); }
template <typename C> LG_CX double cnorm2(C a) {
  // |a|^2 accumulated in FP64 whatever the storage precision (float*float is exact in double)
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// compile-time loop
template <int I, int N, typename F> LG_CX void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(std::forward<F>(f));
  }
}

// cos(2*pi*m/192), m = 0..48, correctly rounded (computed to 40 digits).
constexpr double kQuarterCos192[49] = {
    1.0,
    0.9994645874763656444298364462428599458836,
    0.9978589232386035067380697912727776045318,
    0.9951847266721968862448369531094799215754,
    0.9914448613738104111445575269285628712777,
    0.9866433320848790047469239329842060425036,
    0.9807852804032304491261822361342390369739,
    0.9738769792773336481496899701335503917353,
    0.9659258262890682867497431997288973676339,
    0.9569403357322088649357978869802699694828,
    0.9469301294951056642558042748539836836988,
    0.9359059267573257002917072494667353604861,
    0.9238795325112867561281831893967882868224,
    0.9108638249211758185732917071605506458979,
    0.8968727415326883038941039363930811981791,
    0.8819212643483550297127568636603883495084,
    0.8660254037844386467637231707529361834714,
    0.8492021815265788876490969373431002233934,
    0.8314696123025452370787883776179057567385,
    0.8128466845916152165790961432719088000368,
    0.7933533402912351645797769615012992766286,
    0.7730104533627369608109066097584698009710,
    0.7518398074789773964075194063769614427711,
    0.7298640726978356573501011944031818828671,
    0.7071067811865475244008443621048490392848,
    0.6835923020228712805134975943161551170438,
    0.6593458151000688684251246120553374509154,
    0.6343932841636454982151716132254933706756,
    0.6087614290087206394160975428981640045163,
    0.5824776968678021491971347670361124496849,
    0.5555702330196022247428308139485328743749,
    0.5280678506503679958734488203376055682261,
    0.5,
    0.4713967368259976485563876259052543776574,
    0.4422886902190012819952389773242447301569,
    0.4127070298043947370477021860733674718905,
    0.3826834323650897717284599840303988667613,
    0.3522500479212335065317523197587126718279,
    0.3214394653031615807010576240789015860584,
    0.2902846772544623676361923758173952746914,
    0.2588190451025207623488988376240483283490,
    0.2270762630343732075856966925770880866157,
    0.1950903220161282678482848684770222409276,
    0.1628954733945887394808006397082655595151,
    0.1305261922200515915484062278954890101937,
    0.0980171403295606019941955638886418458611,
    0.0654031292301430668153155587751754414406,
    0.0327190828217761420636599263172881261149,
    0.0};

__host__ __device__ constexpr double cos192(int m) {
  m %= 192;
  if (m < 0) m += 192;
  return m <= 48 ? kQuarterCos192[m] : m <= 96 ? -kQuarterCos192[96 - m] : m <= 144 ? -kQuarterCos192[m - 96]
                                                                                   : kQuarterCos192[192 - m];
}
__host__ __device__ constexpr double sin192(int m) { return cos192(m - 48); }

// a * W_N^E with W_N = exp(-2 pi i / N) (forward sign), N a divisor of 192.
template <int N, int E, typename C> LG_CX C twmul(C a) {
  using R = typename real_of<C>::type;
  static_assert(192 % N == 0, "in-register twiddles support N | 192");
  constexpr int m = ((E % N) * (192 / N)) % 192;
  constexpr R h = (R)kQuarterCos192[24];  // 1/sqrt(2)
  if constexpr (m == 0) return a;
  else if constexpr (m == 48) return mul_mi(a);
  else if constexpr (m == 96) return cneg(a);
  else if constexpr (m == 144) return mul_pi(a);
  else if constexpr (m == 24) return mk<C>((a.x + a.y) * h, (a.y - a.x) * h);
  else if constexpr (m == 72) return mk<C>((a.y - a.x) * h, -(a.x + a.y) * h);
  else if constexpr (m == 120) return mk<C>(-(a.x + a.y) * h, (a.x - a.y) * h);
  else if constexpr (m == 168) return mk<C>((a.x - a.y) * h, (a.x + a.y) * h);
  else {
    constexpr R c = (R)cos192(m);
    constexpr R s = (R)(-sin192(m));
    return mk<C>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// Forward DFT of N points held in registers: X[k] = sum_n x[n] exp(-2 pi i n k / N).
// Natural order in and out; decimation in time.
template <int N, typename C> LG_CX void dft(C (&x)[N]) {
  using R = typename real_of<C>::type;
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
  } else if constexpr (N % 3 == 0) {
    constexpr int M = N / 3;
    constexpr R half_sqrt3 = (R)kQuarterCos192[16];  // cos(pi/6) = sqrt(3)/2
    C s0[M], s1[M], s2[M];
    static_for<0, M>([&](auto mm) {
      constexpr int m = decltype(mm)::value;
      s0[m] = x[3 * m + 0];
      s1[m] = x[3 * m + 1];
      s2[m] = x[3 * m + 2];
    });
    dft<M>(s0);
    dft<M>(s1);
    dft<M>(s2);
    static_for<0, M>([&](auto kk) {
      constexpr int k = decltype(kk)::value;
      C b0 = s0[k];
      C b1 = twmul<N, 1 * k>(s1[k]);
      C b2 = twmul<N, 2 * k>(s2[k]);
      C t1 = cadd(b1, b2), t2 = csub(b1, b2);
      C m1 = mk<C>(b0.x - (R)0.5 * t1.x, b0.y - (R)0.5 * t1.y);
      C m2 = mul_mi(cscale(t2, half_sqrt3));
      x[k] = cadd(b0, t1);
      x[k + M] = cadd(m1, m2);
      x[k + 2 * M] = csub(m1, m2);
    });
  } else {
    static_assert(N % 2 == 0, "dft<N> supports N = 2^a 3^b dividing 192");
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
