// Special functions of the ball-harmonic basis in FP64, for the GPU build of
// the projection tables (and a host build that tests check against SciPy):
//   sph_jn(l, x)            spherical Bessel function of the first kind j_l(x)
//   legendre_on(l, m, x)    orthonormal associated Legendre function with the
//                           Condon-Shortley phase, Y_lm = legendre_on e^{i m phi}
#pragma once

#include <math.h>

namespace mch {
namespace special {

#if defined(__HIPCC__)
#define MCH_SF __host__ __device__ inline
#else
#define MCH_SF inline
#endif

// j_l(x), x >= 0.  Above l the upward recurrence from j_0, j_1 is stable; at
// and below l (and for j_1 below x = 1) the downward (Miller) recurrence from
// well above l and x, rescaled against overflow, normalised by whichever of
// j_0, j_1 is larger in magnitude.
MCH_SF double sph_jn(int l, double x) {
  if (x == 0.0) return l == 0 ? 1.0 : 0.0;
  const double sx = sin(x), cx = cos(x);
  const double j0 = sx / x;
  if (l == 0) return j0;
  const double j1 = sx / (x * x) - cx / x;  // loses digits to cancellation below x = 1
  if (l == 1 && x >= 1.0) return j1;
  if (x > (double)l) {
    double a = j0, b = j1;
    for (int n = 1; n < l; ++n) {
      const double c = (double)(2 * n + 1) / x * b - a;
      a = b;
      b = c;
    }
    return b;
  }
  const int top = (l > (int)x ? l : (int)x) + 20 + (int)sqrt(40.0 * (double)(l + 1));
  double fp1 = 0.0, f = 1e-280, fl = 0.0;
  for (int n = top; n > 0; --n) {
    const double fm1 = (double)(2 * n + 1) / x * f - fp1;  // f_{n-1} from f_n, f_{n+1}
    fp1 = f;
    f = fm1;
    if (n - 1 == l) fl = f;
    if (fabs(f) > 1e250) {
      f *= 1e-250;
      fp1 *= 1e-250;
      fl *= 1e-250;
    }
  }
  // f = f_0, fp1 = f_1 up to one common factor
  return fabs(j0) >= fabs(j1) ? fl * (j0 / f) : fl * (j1 / fp1);
}

// Orthonormal associated Legendre P_l^m(x), m >= 0, with the Condon-Shortley
// phase: P_0^0 = 1 / sqrt(4 pi), P_m^m = -sqrt((2m+1)/(2m)) s P_{m-1}^{m-1}
// (s = sqrt(1 - x^2)), P_{m+1}^m = sqrt(2m+3) x P_m^m, then the three-term
// recurrence in l.
MCH_SF double legendre_on(int l, int m, double x) {
  const double s = sqrt(fmax(0.0, 1.0 - x * x));
  double pmm = 1.0 / sqrt(4.0 * 3.14159265358979323846);
  for (int k = 1; k <= m; ++k) pmm = -sqrt((double)(2 * k + 1) / (2.0 * k)) * s * pmm;
  if (l == m) return pmm;
  double a1 = sqrt((double)(2 * m + 3)) * x * pmm;
  double a2 = pmm;
  for (int ll = m + 2; ll <= l; ++ll) {
    const double a = sqrt((4.0 * ll * ll - 1) / (double)(ll * ll - m * m));
    const double b = sqrt(((ll - 1.0) * (ll - 1.0) - m * m) / (4.0 * (ll - 1) * (ll - 1) - 1));
    const double p = a * (x * a1 - b * a2);
    a2 = a1;
    a1 = p;
  }
  return a1;
}

#undef MCH_SF

}  // namespace special
}  // namespace mch
