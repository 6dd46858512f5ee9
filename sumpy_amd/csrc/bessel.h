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
// bessel_jy_c: J and Y of complex argument by the same route.
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

// 1/a by Smith's method (no |a|^2, so |a| up to the double range: the
// unnormalised Miller sums reach 1e250)
__device__ __forceinline__ cplx cinv(cplx a) {
  if (fabs(a.re) >= fabs(a.im)) {
    const double r = a.im/a.re, d = a.re + a.im*r;
    return {1.0/d, -r/d};
  }
  const double r = a.re/a.im, d = a.re*r + a.im;
  return {r/d, -1.0/d};
}

// J_0..J_n(z) and, when Y != null, Y_0..Y_n(z) for complex z (z != 0 with Y): bessel_jy in complex arithmetic -- Miller's backward
// recurrence normalised with J_0 + 2 sum_k J_2k = 1 and the Neumann series of
// Y_0, Y_1 with the principal branch of log(z/2), Y_n upward.  For moderate
// |Im z| (the normalisation sum holds terms of size e^{|Im z|}).
__device__ void bessel_jy_c(cplx z, int n, cplx *J, cplx *Y) {
  const double az = hypot(z.re, z.im);
  if (az == 0.0) {
    J[0] = {1.0, 0.0};
    for (int m = 1; m <= n; ++m) J[m] = {0.0, 0.0};
    return;
  }
  int M = (n > (int)az ? n : (int)az) + 20 + (int)(16.0*cbrt(az));
  M += M & 1;
  const cplx tz = cscale(cinv(z), 2.0);
  cplx jp1 = {0.0, 0.0}, j = {1.0, 0.0}, norm = {0.0, 0.0}, sy0 = {0.0, 0.0},
       sy1 = {0.0, 0.0};
  for (int m = M; m >= 1; --m) {
    if (m <= n) J[m] = j;
    if ((m & 1) == 0) {
      const int kk = m >> 1;
      norm = cadd(norm, cscale(j, 2.0));
      sy0 = cadd(sy0, cscale(j, ((kk & 1) ? -1.0 : 1.0)/(double)kk));
    } else if (m >= 3) {
      const int kk = (m - 1) >> 1;
      sy1 = cadd(sy1, cscale(j, ((kk & 1) ? -1.0 : 1.0)*(double)(2*kk + 1)/
                                   ((double)kk*(double)(kk + 1))));
    }
    const cplx tj = cmul(tz, j);
    const cplx jm1 = {(double)m*tj.re - jp1.re, (double)m*tj.im - jp1.im};
    jp1 = j;
    j = jm1;
    if (fabs(j.re) + fabs(j.im) > BIG) {
      j = cscale(j, SMALL);
      jp1 = cscale(jp1, SMALL);
      norm = cscale(norm, SMALL);
      sy0 = cscale(sy0, SMALL);
      sy1 = cscale(sy1, SMALL);
      for (int q = m; q <= n; ++q) J[q] = cscale(J[q], SMALL);
    }
  }
  norm = cadd(norm, j);
  J[0] = j;
  const cplx inv = cinv(norm);
  for (int m = 0; m <= n; ++m) J[m] = cmul(J[m], inv);
  if (!Y) return;
  sy0 = cmul(sy0, inv);
  sy1 = cmul(sy1, inv);
  const cplx L = {log(0.5*az) + EULER_GAMMA, atan2(z.im, z.re)};
  const cplx lj0 = cmul(L, J[0]);
  Y[0] = {(2.0/PI)*lj0.re - (4.0/PI)*sy0.re, (2.0/PI)*lj0.im - (4.0/PI)*sy0.im};
  if (n < 1) return;
  const cplx j0z = cmul(J[0], cinv(z));
  const cplx l1j1 = cmul(cplx{L.re - 1.0, L.im}, J[1]);
  Y[1] = {-(2.0/PI)*j0z.re + (2.0/PI)*l1j1.re - (2.0/PI)*sy1.re,
          -(2.0/PI)*j0z.im + (2.0/PI)*l1j1.im - (2.0/PI)*sy1.im};
  for (int m = 1; m < n; ++m) {
    const cplx ty = cmul(tz, Y[m]);
    Y[m + 1] = {(double)m*ty.re - Y[m - 1].re, (double)m*ty.im - Y[m - 1].im};
  }
}

}  // namespace
}  // namespace sumpyhip
