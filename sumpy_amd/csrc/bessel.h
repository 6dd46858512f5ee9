// Bessel function sequences of real argument for the HIP kernels (device
// code, one copy per translation unit).
//
// bessel_jy: J_0..J_n by Miller's backward recurrence normalised with
// J_0 + 2 sum_k J_2k = 1; Y_0 and Y_1 from their Neumann series in the J_n
// (Abramowitz & Stegun 9.1.88), Y_n by the upward recurrence (stable for Y).
// bessel_ik: I_n by Miller's recurrence normalised with I_0 + 2 sum_k I_k =
// e^x; K_0, K_1 by the trapezoidal rule on
// K_nu(x) = int_0^inf e^{-x cosh t} cosh(nu t) dt (step min(1/4, 1/(2 sqrt x)),
// exponentially convergent for this integrand), K_n by the upward recurrence.
#pragma once

#include <cmath>

#include "common.h"

namespace sumpyhip {
namespace {

constexpr double PI = 3.14159265358979323846;
constexpr double EULER_GAMMA = 0.57721566490153286061;
constexpr double BIG = 1e250, SMALL = 1e-250;

// J_0..J_n(x) and, when Y != null, Y_0..Y_n(x) (x > 0 then; n >= 1).
// x = 0: J_0 = 1, J_n = 0 (Y not requested there).
__device__ void bessel_jy(double x, int n, double *J, double *Y) {
  if (x == 0.0) {
    J[0] = 1.0;
    for (int m = 1; m <= n; ++m) J[m] = 0.0;
    return;
  }
  int M = (n > (int)x ? n : (int)x) + 20 + (int)(16.0*cbrt(x));
  M += M & 1;
  const double tox = 2.0/x;
  double jp1 = 0.0, j = 1.0, norm = 0.0, sy0 = 0.0, sy1 = 0.0;
  for (int m = M; m >= 1; --m) {
    // j = J_m, jp1 = J_{m+1} (a common unknown factor)
    if (m <= n) J[m] = j;
    if ((m & 1) == 0) {
      const int kk = m >> 1;
      norm += 2.0*j;
      sy0 += ((kk & 1) ? -j : j)/(double)kk;
    } else if (m >= 3) {
      const int kk = (m - 1) >> 1;
      sy1 += ((kk & 1) ? -j : j)*(double)(2*kk + 1)/((double)kk*(double)(kk + 1));
    }
    const double jm1 = (double)m*tox*j - jp1;
    jp1 = j;
    j = jm1;
    if (fabs(j) > BIG) {
      j *= SMALL; jp1 *= SMALL; norm *= SMALL; sy0 *= SMALL; sy1 *= SMALL;
      for (int q = m; q <= n; ++q) J[q] *= SMALL;
    }
  }
  norm += j;
  J[0] = j;
  const double inv = 1.0/norm;
  for (int m = 0; m <= n; ++m) J[m] *= inv;
  if (!Y) return;
  sy0 *= inv;
  sy1 *= inv;
  const double L = log(0.5*x) + EULER_GAMMA;
  Y[0] = (2.0/PI)*L*J[0] - (4.0/PI)*sy0;
  Y[1] = -2.0*J[0]/(PI*x) + (2.0/PI)*(L - 1.0)*J[1] - (2.0/PI)*sy1;
  for (int m = 1; m < n; ++m) Y[m + 1] = (double)m*tox*Y[m] - Y[m - 1];
}

// I_0..I_n(x) and, when K != null, K_0..K_n(x) (x > 0; n >= 1 with K).
// x = 0: I_0 = 1, I_n = 0 (K not requested there).
__device__ void bessel_ik(double x, int n, double *I, double *K) {
  if (x == 0.0) {
    I[0] = 1.0;
    for (int m = 1; m <= n; ++m) I[m] = 0.0;
    return;
  }
  int M = (n > (int)x ? n : (int)x) + 20 + (int)(16.0*cbrt(x));
  M += M & 1;
  const double tox = 2.0/x;
  double ip1 = 0.0, v = 1.0, norm = 0.0;
  for (int m = M; m >= 1; --m) {
    if (m <= n) I[m] = v;
    norm += 2.0*v;
    const double im1 = (double)m*tox*v + ip1;
    ip1 = v;
    v = im1;
    if (v > BIG) {
      v *= SMALL; ip1 *= SMALL; norm *= SMALL;
      for (int q = m; q <= n; ++q) I[q] *= SMALL;
    }
  }
  norm += v;
  I[0] = v;
  const double f = exp(x)/norm;
  for (int m = 0; m <= n; ++m) I[m] *= f;
  if (!K) return;
  const double h = x <= 4.0 ? 0.25 : 0.5/sqrt(x);
  const double e0 = exp(-x);
  double s0 = 0.5*e0, s1 = 0.5*e0;
  for (int j = 1; j < 4096; ++j) {
    const double ct = cosh(j*h), e = exp(-x*ct);
    s0 += e;
    s1 += e*ct;
    if (x*(ct - 1.0) > 40.0) break;
  }
  K[0] = h*s0;
  K[1] = h*s1;
  for (int m = 1; m < n; ++m) K[m + 1] = (double)m*tox*K[m] + K[m - 1];
}

}  // namespace
}  // namespace sumpyhip
