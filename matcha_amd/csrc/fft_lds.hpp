// Compile-time four-step FFTs for lines held in LDS (gfx950).
//
// A length N = N1 * N2 transform is done in two in-place passes over LDS:
//   forward (natural order in, "transposed" order out)
//     step 1: for each n2, gather x[N2 n1 + n2] (n1 < N1), N1-point DFT in
//             registers, multiply by W_N^{n2 k1}, store at slot N2 k1 + n2
//     step 2: for each k1, N2-point DFT of slots k1 N2 + (0..N2-1), store in
//             place: slot k1 N2 + k2 holds X[k1 + N1 k2]
//   inverse (transposed order in, natural order out)
//     step A: for each k1, N2-point DFT of slots k1 N2 + k2, multiply by
//             W_N^{k1 n2}, store in place at slot k1 N2 + n2
//     step B: for each n2, N1-point DFT of slots k1 N2 + n2, store x[n2 + N2 n1]
// Every work item reads and writes the same set of slots, so a step needs no
// barrier between its loads and stores, only one after it.  Consumers of the
// forward output (pointwise products, argmax) map slots to frequencies with
// freq_of_slot().  The register DFTs are radix-4/3/2 compositions with
// compile-time twiddles (odd primes 5, 7 by their symmetric pairs); W_N^{j}
// comes from an LDS table tw[j] = e^{-2 pi i j/N}.
//
// Work items are laid out with lanes along lines (consecutive lanes handle
// consecutive lines), which is conflict-free in LDS when lines are contiguous
// rows of odd stride or columns of a row-major tile.
//
// The step functions are __host__ __device__ and take the thread index, so
// that a host program can run them thread by thread between "barriers".
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace mch {
namespace fft {

#define MCH_HD __host__ __device__ __forceinline__

constexpr double kPi = 3.14159265358979323846264338327950288;

// Taylor series on [-pi, pi]; exact to double rounding for these arguments.
constexpr double csin(double x) {
  double t = x, s = x;
  for (int k = 1; k < 24; ++k) {
    t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}
constexpr double ccos(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 24; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}
// angle of e^{sign 2 pi i j / R} reduced to [-pi, pi]
constexpr double tw_angle(int j, int R, int sign) {
  const int jj = ((j % R) + R) % R;
  double a = 2.0 * kPi * jj / R;
  if (a > kPi) a -= 2.0 * kPi;
  return sign * a;
}

MCH_HD float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
MCH_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
MCH_HD float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (i c)
MCH_HD float2 cmuli(float2 a, float c) { return make_float2(-c * a.y, c * a.x); }

// v * e^{SIGN 2 pi i J / R}
template <int J, int R, int SIGN>
MCH_HD float2 twiddle(float2 v) {
  constexpr int j = ((J % R) + R) % R;
  if constexpr (j == 0) {
    return v;
  } else if constexpr (2 * j == R) {
    return make_float2(-v.x, -v.y);
  } else if constexpr (4 * j == R) {  // SIGN * i
    return SIGN > 0 ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
  } else if constexpr (4 * j == 3 * R) {  // -SIGN * i
    return SIGN > 0 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
  } else {
    constexpr double a = tw_angle(j, R, SIGN);
    constexpr float c = (float)ccos(a);
    constexpr float s = (float)csin(a);
    return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

template <class F, int... I>
MCH_HD void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
MCH_HD void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// the first factor of a composite register DFT: 4, then 3, then 2
constexpr int dft_split(int R) { return R % 4 == 0 ? 4 : (R % 3 == 0 ? 3 : (R % 2 == 0 ? 2 : 1)); }

// In-register DFT of R points, natural order in and out:
// X[k] = sum_n v[n] e^{SIGN 2 pi i n k / R}.  Composite R = A B (A = 4, 3 or
// 2) as A-point DFTs, twiddles and B-point DFTs; R = 2, 3, 4, 5, 7 directly.
template <int R, int SIGN>
struct Dft {
  static constexpr int A = dft_split(R);
  static constexpr int B = R / A;
  static_assert(A > 1 && B > 1, "register DFT sizes: products of 2, 3, 4, 5 and 7");
  static MCH_HD void run(float2* v) {
    float2 t[R];
    static_for<B>([&](auto n2c) {
      constexpr int n2 = decltype(n2c)::value;
      float2 u[A];
      static_for<A>([&](auto n1c) {
        constexpr int n1 = decltype(n1c)::value;
        u[n1] = v[B * n1 + n2];
      });
      Dft<A, SIGN>::run(u);
      static_for<A>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        t[k1 * B + n2] = twiddle<n2 * k1, R, SIGN>(u[k1]);
      });
    });
    static_for<A>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      float2 u[B];
      static_for<B>([&](auto n2c) {
        constexpr int n2 = decltype(n2c)::value;
        u[n2] = t[k1 * B + n2];
      });
      Dft<B, SIGN>::run(u);
      static_for<B>([&](auto k2c) {
        constexpr int k2 = decltype(k2c)::value;
        v[k1 + A * k2] = u[k2];
      });
    });
  }
};

// Odd prime R: X[k] = v0 + sum_{n <= (R-1)/2} (v_n + v_{R-n}) cos(2 pi n k / R)
//                     + SIGN i (v_n - v_{R-n}) sin(2 pi n k / R).
template <int R, int SIGN>
struct DftOdd {
  static MCH_HD void run(float2* v) {
    constexpr int H = (R - 1) / 2;
    float2 s[H + 1], d[H + 1];
    static_for<H>([&](auto nc) {
      constexpr int n = decltype(nc)::value + 1;
      s[n] = cadd(v[n], v[R - n]);
      d[n] = csub(v[n], v[R - n]);
    });
    float2 out[R];
    float2 x0 = v[0];
    static_for<H>([&](auto nc) { x0 = cadd(x0, s[decltype(nc)::value + 1]); });
    out[0] = x0;
    static_for<H>([&](auto kc) {
      constexpr int k = decltype(kc)::value + 1;
      float2 re = v[0], im = make_float2(0.f, 0.f);
      static_for<H>([&](auto nc) {
        constexpr int n = decltype(nc)::value + 1;
        constexpr double a = tw_angle(n * k, R, 1);
        constexpr float c = (float)ccos(a), sn = (float)csin(a);
        re = make_float2(re.x + c * s[n].x, re.y + c * s[n].y);
        im = make_float2(im.x + sn * d[n].x, im.y + sn * d[n].y);
      });
      const float2 q = cmuli(im, (float)SIGN);  // SIGN i im
      out[k] = cadd(re, q);
      out[R - k] = csub(re, q);
    });
    static_for<R>([&](auto kc) { v[decltype(kc)::value] = out[decltype(kc)::value]; });
  }
};
template <int SIGN> struct Dft<5, SIGN> : DftOdd<5, SIGN> {};
template <int SIGN> struct Dft<7, SIGN> : DftOdd<7, SIGN> {};

template <int SIGN>
struct Dft<2, SIGN> {
  static MCH_HD void run(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int SIGN>
struct Dft<3, SIGN> {
  static MCH_HD void run(float2* v) {
    const float2 t = cadd(v[1], v[2]);
    const float2 d = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    const float2 q = cmuli(d, (float)SIGN * 0.86602540378443864676f);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, q);
    v[2] = csub(m, q);
  }
};

template <int SIGN>
struct Dft<4, SIGN> {
  static MCH_HD void run(float2* v) {
    const float2 a = cadd(v[0], v[2]);
    const float2 b = csub(v[0], v[2]);
    const float2 c = cadd(v[1], v[3]);
    const float2 d = cmuli(csub(v[1], v[3]), (float)SIGN);
    v[0] = cadd(a, c);
    v[1] = cadd(b, d);
    v[2] = csub(a, c);
    v[3] = csub(b, d);
  }
};

// N = N1 * N2 (supported line lengths; the factors are register DFT sizes)
template <int N>
struct Shape;
template <> struct Shape<16> { static constexpr int N1 = 4, N2 = 4; };
template <> struct Shape<32> { static constexpr int N1 = 4, N2 = 8; };
template <> struct Shape<48> { static constexpr int N1 = 4, N2 = 12; };
template <> struct Shape<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Shape<96> { static constexpr int N1 = 8, N2 = 12; };
template <> struct Shape<128> { static constexpr int N1 = 8, N2 = 16; };
template <> struct Shape<80> { static constexpr int N1 = 8, N2 = 10; };
template <> struct Shape<112> { static constexpr int N1 = 7, N2 = 16; };
template <> struct Shape<144> { static constexpr int N1 = 12, N2 = 12; };
template <> struct Shape<160> { static constexpr int N1 = 10, N2 = 16; };
template <> struct Shape<192> { static constexpr int N1 = 12, N2 = 16; };

// frequency held by slot s of a forward (transposed-order) output, and back
template <int N>
MCH_HD int freq_of_slot(int s) {
  return s / Shape<N>::N2 + Shape<N>::N1 * (s % Shape<N>::N2);
}
template <int N>
MCH_HD int slot_of_freq(int f) {
  return (f % Shape<N>::N1) * Shape<N>::N2 + f / Shape<N>::N1;
}

template <int SIGN>
MCH_HD float2 table_twiddle(const float2* tw, int j) {
  float2 w = tw[j];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// NL lines: element n of line l at buf[l * ls + n * es]; NT threads.
template <int N, int NL, int NT, int SIGN>
MCH_HD void fwd_step1(float2* buf, int ls, int es, const float2* tw, int tid) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, ITEMS = NL * N2;
#pragma unroll
  for (int it = 0; it < (ITEMS + NT - 1) / NT; ++it) {
    const int item = it * NT + tid;
    if (ITEMS % NT != 0 && item >= ITEMS) break;
    const int line = item % NL, n2 = item / NL;
    float2* p = buf + line * ls + n2 * es;
    float2 v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = p[n1 * N2 * es];
    Dft<N1, SIGN>::run(v);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], table_twiddle<SIGN>(tw, n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) p[k1 * N2 * es] = v[k1];
  }
}

template <int N, int NL, int NT, int SIGN>
MCH_HD void fwd_step2(float2* buf, int ls, int es, int tid) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, ITEMS = NL * N1;
#pragma unroll
  for (int it = 0; it < (ITEMS + NT - 1) / NT; ++it) {
    const int item = it * NT + tid;
    if (ITEMS % NT != 0 && item >= ITEMS) break;
    const int line = item % NL, k1 = item / NL;
    float2* p = buf + line * ls + k1 * N2 * es;
    float2 u[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) u[n2] = p[n2 * es];
    Dft<N2, SIGN>::run(u);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) p[k2 * es] = u[k2];
  }
}

template <int N, int NL, int NT, int SIGN>
MCH_HD void inv_stepA(float2* buf, int ls, int es, const float2* tw, int tid) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, ITEMS = NL * N1;
#pragma unroll
  for (int it = 0; it < (ITEMS + NT - 1) / NT; ++it) {
    const int item = it * NT + tid;
    if (ITEMS % NT != 0 && item >= ITEMS) break;
    const int line = item % NL, k1 = item / NL;
    float2* p = buf + line * ls + k1 * N2 * es;
    float2 u[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) u[k2] = p[k2 * es];
    Dft<N2, SIGN>::run(u);
#pragma unroll
    for (int n2 = 1; n2 < N2; ++n2) u[n2] = cmul(u[n2], table_twiddle<SIGN>(tw, k1 * n2));
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) p[n2 * es] = u[n2];
  }
}

template <int N, int NL, int NT, int SIGN>
MCH_HD void inv_stepB(float2* buf, int ls, int es, int tid) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, ITEMS = NL * N2;
#pragma unroll
  for (int it = 0; it < (ITEMS + NT - 1) / NT; ++it) {
    const int item = it * NT + tid;
    if (ITEMS % NT != 0 && item >= ITEMS) break;
    const int line = item % NL, n2 = item / NL;
    float2* p = buf + line * ls + n2 * es;
    float2 v[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[k1] = p[k1 * N2 * es];
    Dft<N1, SIGN>::run(v);
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) p[n1 * N2 * es] = v[n1];
  }
}

#undef MCH_HD

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// Whole transforms for a workgroup of NT threads (ends with a barrier).
template <int N, int NL, int NT, int SIGN>
__device__ __forceinline__ void forward(float2* buf, int ls, int es, const float2* tw) {
  fwd_step1<N, NL, NT, SIGN>(buf, ls, es, tw, threadIdx.x);
  __syncthreads();
  fwd_step2<N, NL, NT, SIGN>(buf, ls, es, threadIdx.x);
  __syncthreads();
}
template <int N, int NL, int NT, int SIGN>
__device__ __forceinline__ void inverse(float2* buf, int ls, int es, const float2* tw) {
  inv_stepA<N, NL, NT, SIGN>(buf, ls, es, tw, threadIdx.x);
  __syncthreads();
  inv_stepB<N, NL, NT, SIGN>(buf, ls, es, threadIdx.x);
  __syncthreads();
}
#endif

}  // namespace fft
}  // namespace mch
