// Host check of the LDS four-step FFTs (matcha_amd/csrc/fft_lds.hpp): every
// supported line length, forward and inverse, run thread by thread on the CPU
// with a "barrier" between the steps, against a direct DFT in double.
// Prints "N max_rel_err_fwd max_rel_err_roundtrip" per length.
#include <cmath>
#include <cstdio>
#include <vector>

#include "fft_lds.hpp"

using namespace mch::fft;

template <int N>
void check() {
  constexpr int NL = 3, NT = 64;
  std::vector<float2> tw(N);
  for (int j = 0; j < N; ++j) tw[j] = make_float2((float)cos(-2 * M_PI * j / N), (float)sin(-2 * M_PI * j / N));
  std::vector<float2> buf(NL * N), orig(NL * N);
  unsigned s = 12345u + N;
  for (auto& v : orig) {
    s = s * 1664525u + 1013904223u;
    const float a = (float)((s >> 8) & 0xffff) / 65536.f - 0.5f;
    s = s * 1664525u + 1013904223u;
    const float b = (float)((s >> 8) & 0xffff) / 65536.f - 0.5f;
    v = make_float2(a, b);
  }
  buf = orig;
  // forward, SIGN = -1: X[k] = sum_n x[n] e^{-2 pi i n k / N}, slots in transposed order
  for (int t = 0; t < NT; ++t) fwd_step1<N, NL, NT, -1>(buf.data(), N, 1, tw.data(), t);
  for (int t = 0; t < NT; ++t) fwd_step2<N, NL, NT, -1>(buf.data(), N, 1, t);
  double err_f = 0.0, scale = 0.0;
  for (int l = 0; l < NL; ++l)
    for (int k = 0; k < N; ++k) {
      double re = 0, im = 0;
      for (int n = 0; n < N; ++n) {
        const double a = -2 * M_PI * (double)n * k / N;
        re += orig[l * N + n].x * cos(a) - orig[l * N + n].y * sin(a);
        im += orig[l * N + n].x * sin(a) + orig[l * N + n].y * cos(a);
      }
      const float2 g = buf[l * N + slot_of_freq<N>(k)];
      err_f = fmax(err_f, hypot(g.x - re, g.y - im));
      scale = fmax(scale, hypot(re, im));
      if (freq_of_slot<N>(slot_of_freq<N>(k)) != k) err_f = 1e9;
    }
  // inverse, SIGN = +1, back to natural order: N x
  for (int t = 0; t < NT; ++t) inv_stepA<N, NL, NT, 1>(buf.data(), N, 1, tw.data(), t);
  for (int t = 0; t < NT; ++t) inv_stepB<N, NL, NT, 1>(buf.data(), N, 1, t);
  double err_r = 0.0;
  for (int i = 0; i < NL * N; ++i)
    err_r = fmax(err_r, hypot(buf[i].x / N - orig[i].x, buf[i].y / N - orig[i].y));
  printf("%d %.3e %.3e\n", N, err_f / scale, err_r);
}

int main() {
  check<16>(); check<32>(); check<48>(); check<64>(); check<80>(); check<96>();
  check<112>(); check<128>(); check<144>(); check<160>(); check<192>();
  return 0;
}
