// Host build of the FP64 special functions of the GPU basis build
// (matcha_amd/csrc/special.hpp).  Prints "j l x value" for a grid of orders
// and arguments, and "p l m x value" for the orthonormal Legendre functions.
#include <cstdio>

#include "special.hpp"

int main() {
  const double xs[] = {0.0, 1e-3, 0.05, 0.7, 1.0, 2.5, 3.14159, 7.9, 12.0, 25.0, 33.3, 49.0, 60.5, 75.0, 101.0, 140.0};
  for (int l = 0; l <= 100; l += (l < 10 ? 1 : 3))
    for (double x : xs) std::printf("j %d %.17g %.17g\n", l, x, mch::special::sph_jn(l, x));
  const double cs[] = {-1.0, -0.93, -0.5, -0.1, 0.0, 0.2, 0.61, 0.999, 1.0};
  for (int l = 0; l <= 90; l += (l < 8 ? 1 : 7))
    for (int m = 0; m <= l; m += (m < 4 ? 1 : 5))
      for (double x : cs) std::printf("p %d %d %.17g %.17g\n", l, m, x, mch::special::legendre_on(l, m, x));
  return 0;
}
