// Line-Taylor QBX for the Laplace, Helmholtz and Yukawa kernels in 2D and 3D.
//
// Reference: LayerPotential (sumpy/qbx.py:250-335) with the
// LineTaylorLocalExpansion (sumpy/expansion/local.py:116-198): per (target t,
// source s) with centre c, a = c - s, b = t - c, the expansion coefficients
// are coeff_i = D_s d^i/dtau^i K(a + tau b) at tau = 0 and the value is
// sum_{i<=p} coeff_i / i!, i.e. the order-p Taylor polynomial of
// tau -> K(a + tau b) evaluated at tau = 1.  D_s is the source derivative of
// the source kernel (-n . grad_a for DirectionalSourceDerivative,
// -d/da_j for AxisSourceDerivative; sumpy/kernel.py:2457-2507); a target
// derivative AxisTargetDerivative differentiates the coefficients in a_j
// (LayerPotentialBase._evaluate, sumpy/qbx.py:125-147).
//
// Kernels (expression without the scaling constant, sumpy/kernel.py:639-897):
//   0 Laplace 2D  log r        1 Laplace 3D  1/r
//   2 Helmholtz 2D H0(k r)     3 Helmholtz 3D e^{ikr}/r
//   4 Yukawa 2D   H0(i lam r)  5 Yukawa 3D   e^{-lam r}/r
//
// Method (power-series arithmetic, no symbolic differentiation): with
// rho(tau) = |a + tau b|^2 = A + B tau + C tau^2, the series of
// r(tau) = sqrt(rho(tau)) follows from R_0 = sqrt(A),
// R_i = (rho_i - sum_{j=1}^{i-1} R_j R_{i-j}) / (2 R_0).  The radial function
// and its first two derivatives are composed with r(tau) = r0 + delta(tau)
// through their Taylor coefficients g_n = G^(n)(r0)/n! at r0 = |a|:
// G(r(tau)) = sum_n g_n delta(tau)^n (delta has no constant term, so order p
// needs n <= p).  The vector factors are then exact polynomials in tau:
//   d(tau) = a + tau b,   grad K = d G'(r)/r,
//   DLP   -n.grad K = -(n.d) G'/r,
//   d/da_j of that = -n_j G'/r - (n.d) d_j (G'' - G'/r)/r^2,
// with the series divisions by r(tau) done by the usual recurrence.  Radial
// Taylor coefficients: log r and 1/r in closed form; e^{s r}/r as the Cauchy
// product of e^{s r0} s^m/m! and (-1)^j/r0^{j+1}; H0(s r) from
// d^n/dz^n H0(z) = 2^-n sum_j (-1)^(n-j) C(n,j) H_{n-2j}(z) (H_{-m} = (-1)^m H_m)
// with the Bessel sequences of bessel.h.
//
// Layout: one thread per target (apply) or per matrix entry; the sources of
// the apply kernel are staged in LDS in tiles of 256.  Values are complex
// (re, im interleaved) in FP64.  Roofline: VALU (O(p^2) complex operations
// per pair).
#include "common.h"
#include "bessel.h"

namespace sumpyhip {
namespace {

constexpr int QMAX = 12;        // highest supported order
constexpr int QS = QMAX + 3;    // series / Taylor lengths (order + 2 derivatives)

enum { LAP2 = 0, LAP3 = 1, HELM2 = 2, HELM3 = 3, YUK2 = 4, YUK3 = 5 };

template <int KK>
struct KDim {
  static constexpr int value = (KK == LAP3 || KK == HELM3 || KK == YUK3) ? 3 : 2;
};

__device__ __forceinline__ cplx cfma_s(cplx a, double s, cplx acc) {
  return {acc.re + a.re*s, acc.im + a.im*s};
}

// g[n] = G^(n)(r0)/n!, n = 0..ng, of the radial expression G; kk = k
// (Helmholtz, complex allowed in 3D) or (lam, 0) (Yukawa).
template <int KK>
__device__ void radial_taylor(double r0, cplx kk, int ng, cplx *g) {
  const double ip = 1.0/r0;
  if (KK == LAP2) {
    g[0] = {log(r0), 0.0};
    double pw = ip;
    for (int n = 1; n <= ng; ++n) {
      g[n] = {((n & 1) ? pw : -pw)/n, 0.0};
      pw *= ip;
    }
  } else if (KK == LAP3) {
    double pw = ip;
    for (int n = 0; n <= ng; ++n) {
      g[n] = {(n & 1) ? -pw : pw, 0.0};
      pw *= ip;
    }
  } else if (KK == HELM3 || KK == YUK3) {
    // s = i k or -lam
    const cplx s = KK == HELM3 ? cplx{-kk.im, kk.re} : cplx{-kk.re, 0.0};
    const double er = exp(s.re*r0);
    const cplx e = {er*cos(s.im*r0), er*sin(s.im*r0)};
    cplx em[QS];
    double inv[QS];
    em[0] = {1.0, 0.0};
    inv[0] = ip;
    for (int m = 1; m <= ng; ++m) {
      em[m] = cscale(cmul(em[m - 1], s), 1.0/m);
      inv[m] = -inv[m - 1]*ip;
    }
    for (int n = 0; n <= ng; ++n) {
      cplx acc = {0.0, 0.0};
      for (int m = 0; m <= n; ++m) acc = cfma_s(em[m], inv[n - m], acc);
      g[n] = cmul(e, acc);
    }
  } else {
    // H0(s r), s = k (real or complex) or i lam; Z[m] = H_m(s r0)
    double A[QS + 1], Bv[QS + 1];
    cplx Z[QS + 1];
    const int nz = ng < 1 ? 1 : ng;
    if (KK == HELM2 && kk.im != 0.0) {
      cplx Jc[QS + 1], Yc[QS + 1];
      bessel_jy_c(cscale(kk, r0), nz, Jc, Yc);
      for (int m = 0; m <= nz; ++m) Z[m] = {Jc[m].re - Yc[m].im, Jc[m].im + Yc[m].re};
    } else if (KK == HELM2) {
      bessel_jy(kk.re*r0, nz, A, Bv);
      for (int m = 0; m <= nz; ++m) Z[m] = {A[m], Bv[m]};
    } else {
      bessel_ik(kk.re*r0, nz, A, Bv);
      for (int m = 0; m <= nz; ++m) {
        const double v = (2.0/PI)*Bv[m];  // (2/pi) i^(-m-1) K_m
        const int q = m & 3;
        Z[m] = q == 0 ? cplx{0.0, -v} : q == 1 ? cplx{-v, 0.0} : q == 2 ? cplx{0.0, v}
                                                                        : cplx{v, 0.0};
      }
    }
    // s^n / (2^n n!)
    cplx sp = {1.0, 0.0};
    const cplx s = KK == HELM2 ? kk : cplx{0.0, kk.re};
    for (int n = 0; n <= ng; ++n) {
      cplx acc = {0.0, 0.0};
      double binom = 1.0;  // C(n, j)
      for (int j = 0; j <= n; ++j) {
        const int m = n - 2*j;
        const cplx h = m >= 0 ? Z[m] : ((-m) & 1 ? cplx{-Z[-m].re, -Z[-m].im} : Z[-m]);
        acc = cfma_s(h, ((n - j) & 1) ? -binom : binom, acc);
        binom = binom*(n - j)/(j + 1);
      }
      g[n] = cmul(sp, acc);
      sp = cscale(cmul(sp, s), 0.5/(n + 1));
    }
  }
}

// out = num / R as series to order p (R real, R[0] != 0)
__device__ __forceinline__ void series_div(int p, const cplx *num, const double *R, cplx *out) {
  const double i0 = 1.0/R[0];
  for (int i = 0; i <= p; ++i) {
    cplx acc = num[i];
    for (int j = 1; j <= i; ++j) acc = cfma_s(out[i - j], -R[j], acc);
    out[i] = cscale(acc, i0);
  }
}

// sum_{i<=p} [tau^i] (al + tau be) S(tau)
__device__ __forceinline__ cplx lin_sum(int p, double al, double be, const cplx *S) {
  cplx acc = {0.0, 0.0};
  for (int i = 0; i <= p; ++i) acc = cfma_s(S[i], al, acc);
  for (int i = 0; i < p; ++i) acc = cfma_s(S[i], be, acc);
  return acc;
}

// The line-Taylor value of one (target, source) pair (without the scaling
// constant and the density).  DLP: source derivative along nrm; TD: target
// derivative along axis tax.
template <int KK, bool DLP, bool TD>
__device__ cplx line_value(int p, const double *a, const double *b, const double *nrm, int tax,
                           cplx kk) {
  constexpr int D = KDim<KK>::value;
  double A = 0.0, B = 0.0, C = 0.0;
  for (int d = 0; d < D; ++d) {
    A += a[d]*a[d];
    B += 2.0*a[d]*b[d];
    C += b[d]*b[d];
  }
  double R[QS];
  R[0] = sqrt(A);
  const double i2r = 0.5/R[0];
  for (int i = 1; i <= p; ++i) {
    double acc = i == 1 ? B : (i == 2 ? C : 0.0);
    for (int j = 1; j < i; ++j) acc -= R[j]*R[i - j];
    R[i] = acc*i2r;
  }
  constexpr int NDER = DLP ? (TD ? 2 : 1) : (TD ? 1 : 0);
  cplx g[QS + 2];
  radial_taylor<KK>(R[0], kk, p + NDER, g);
  // F0 = G(r(tau)), F1 = G'(r(tau)), F2 = G''(r(tau)) by powers of delta
  cplx F0[QS], F1[QS], F2[QS];
  double Dp[QS];
  for (int i = 0; i <= p; ++i) {
    F0[i] = F1[i] = F2[i] = {0.0, 0.0};
    Dp[i] = i == 0 ? 1.0 : 0.0;
  }
  for (int n = 0; n <= p; ++n) {
    const cplx c1 = cscale(g[n + 1 <= p + NDER ? n + 1 : 0], (double)(n + 1));
    const cplx c2 = cscale(g[n + 2 <= p + NDER ? n + 2 : 0], (double)((n + 1)*(n + 2)));
    for (int i = n; i <= p; ++i) {
      if (NDER == 0) F0[i] = cfma_s(g[n], Dp[i], F0[i]);
      if (NDER >= 1) F1[i] = cfma_s(c1, Dp[i], F1[i]);
      if (NDER == 2) F2[i] = cfma_s(c2, Dp[i], F2[i]);
    }
    // Dp *= delta (delta_0 = 0, delta_j = R_j)
    for (int i = p; i >= 0; --i) {
      double acc = 0.0;
      for (int j = 1; j <= i; ++j) acc += R[j]*Dp[i - j];
      Dp[i] = acc;
    }
  }
  if (NDER == 0) {
    cplx acc = {0.0, 0.0};
    for (int i = 0; i <= p; ++i) acc = cadd(acc, F0[i]);
    return acc;
  }
  cplx H1[QS];
  series_div(p, F1, R, H1);  // G'(r)/r
  double na = 0.0, nb = 0.0;
  if (DLP)
    for (int d = 0; d < D; ++d) {
      na += nrm[d]*a[d];
      nb += nrm[d]*b[d];
    }
  if (!DLP) return lin_sum(p, a[tax], b[tax], H1);
  if (!TD) {
    const cplx v = lin_sum(p, na, nb, H1);
    return {-v.re, -v.im};
  }
  // -n_j G'/r - (n.d)(d_j) (G'' - G'/r)/r^2
  cplx E[QS], T[QS];
  for (int i = 0; i <= p; ++i) E[i] = {F2[i].re - H1[i].re, F2[i].im - H1[i].im};
  series_div(p, E, R, T);
  series_div(p, T, R, E);
  // (na + tau nb)(a_j + tau b_j) = q0 + q1 tau + q2 tau^2
  const double q0 = na*a[tax], q1 = na*b[tax] + nb*a[tax], q2 = nb*b[tax];
  cplx acc = {0.0, 0.0};
  for (int i = 0; i <= p; ++i) {
    acc = cfma_s(H1[i], -nrm[tax], acc);
    acc = cfma_s(E[i], -q0, acc);
    if (i >= 1) acc = cfma_s(E[i - 1], -q1, acc);
    if (i >= 2) acc = cfma_s(E[i - 2], -q2, acc);
  }
  return acc;
}

struct LineArgs {
  int p, src_kind, src_axis, tgt_axis;
  cplx kk, scale;
  int nt, ns;
  const double *t[3], *c[3], *s[3], *n[3];
  const double *density;
  int accumulate;
  double *out;
};

constexpr int TILE = 256;

template <int KK, bool DLP, bool TD>
__global__ __launch_bounds__(TILE) void qbx_line_apply(LineArgs g) {
  constexpr int D = KDim<KK>::value;
  __shared__ double sS[TILE][2*D + 1];
  const int it = blockIdx.x*TILE + threadIdx.x;
  const bool act = it < g.nt;
  double a0[D], b[D];
  for (int d = 0; d < D; ++d) {
    const double cd = act ? g.c[d][it] : 0.0;
    a0[d] = cd;
    b[d] = act ? g.t[d][it] - cd : 0.0;
  }
  cplx acc = {0.0, 0.0};
  for (int s0 = 0; s0 < g.ns; s0 += TILE) {
    __syncthreads();
    const int js = s0 + threadIdx.x;
    if (js < g.ns) {
      for (int d = 0; d < D; ++d) {
        sS[threadIdx.x][d] = g.s[d][js];
        double nv = 0.0;
        if (DLP) nv = g.src_kind == 2 ? (d == g.src_axis ? 1.0 : 0.0) : g.n[d][js];
        sS[threadIdx.x][D + d] = nv;
      }
      sS[threadIdx.x][2*D] = g.density[js];
    }
    __syncthreads();
    const int cnt = min(TILE, g.ns - s0);
    if (!act) continue;
    for (int j = 0; j < cnt; ++j) {
      double a[D], nrm[D];
      for (int d = 0; d < D; ++d) {
        a[d] = a0[d] - sS[j][d];
        nrm[d] = sS[j][D + d];
      }
      const cplx v = line_value<KK, DLP, TD>(g.p, a, b, nrm, g.tgt_axis, g.kk);
      acc = cfma_s(v, sS[j][2*D], acc);
    }
  }
  if (!act) return;
  const cplx r = cmul(acc, g.scale);
  double *o = g.out + 2*(size_t)it;
  if (g.accumulate) {
    o[0] += r.re;
    o[1] += r.im;
  } else {
    o[0] = r.re;
    o[1] = r.im;
  }
}

template <int KK, bool DLP, bool TD>
__global__ __launch_bounds__(TILE) void qbx_line_entries(LineArgs g, long long nent,
                                                        const int *ti, const int *si) {
  constexpr int D = KDim<KK>::value;
  const size_t e = blockIdx.x*(size_t)TILE + threadIdx.x;
  if (e >= (size_t)nent) return;
  // without index arrays: all nt x ns entries, row-major
  const int it = ti ? ti[e] : (int)(e/(size_t)g.ns);
  const int js = si ? si[e] : (int)(e - (size_t)it*g.ns);
  double a[D], b[D], nrm[D];
  for (int d = 0; d < D; ++d) {
    const double cd = g.c[d][it];
    a[d] = cd - g.s[d][js];
    b[d] = g.t[d][it] - cd;
    nrm[d] = 0.0;
    if (DLP) nrm[d] = g.src_kind == 2 ? (d == g.src_axis ? 1.0 : 0.0) : g.n[d][js];
  }
  const cplx r = cmul(line_value<KK, DLP, TD>(g.p, a, b, nrm, g.tgt_axis, g.kk), g.scale);
  double *o = g.out + 2*e;
  if (g.accumulate) {
    o[0] += r.re;
    o[1] += r.im;
  } else {
    o[0] = r.re;
    o[1] = r.im;
  }
}

template <int KK, bool DLP, bool TD>
int launch_line(const LineArgs &g, long long nent, const int *ti, const int *si, hipStream_t st) {
  if (nent < 0) {
    qbx_line_apply<KK, DLP, TD><<<(g.nt + TILE - 1)/TILE, TILE, 0, st>>>(g);
  } else {
    if (nent == 0) return SUMPYHIP_OK;
    qbx_line_entries<KK, DLP, TD><<<(unsigned)((nent + TILE - 1)/TILE), TILE, 0, st>>>(g, nent, ti,
                                                                                       si);
  }
  return check_launch("qbx_line");
}

template <int KK>
int dispatch_line_kk(const LineArgs &g, long long nent, const int *ti, const int *si,
                     hipStream_t st) {
  const bool dlp = g.src_kind != 0, td = g.tgt_axis >= 0;
  if (dlp)
    return td ? launch_line<KK, true, true>(g, nent, ti, si, st)
              : launch_line<KK, true, false>(g, nent, ti, si, st);
  return td ? launch_line<KK, false, true>(g, nent, ti, si, st)
            : launch_line<KK, false, false>(g, nent, ti, si, st);
}

int dispatch_line(int kind, const LineArgs &g, long long nent, const int *ti, const int *si,
                  hipStream_t st) {
  switch (kind) {
    case LAP2: return dispatch_line_kk<LAP2>(g, nent, ti, si, st);
    case LAP3: return dispatch_line_kk<LAP3>(g, nent, ti, si, st);
    case HELM2: return dispatch_line_kk<HELM2>(g, nent, ti, si, st);
    case HELM3: return dispatch_line_kk<HELM3>(g, nent, ti, si, st);
    case YUK2: return dispatch_line_kk<YUK2>(g, nent, ti, si, st);
    default: return dispatch_line_kk<YUK3>(g, nent, ti, si, st);
  }
}

int fill_line_args(LineArgs &g, int kernel_kind, int order, int src_kind, int src_axis,
                   int tgt_axis, double k_re, double k_im, int ntargets, const double *tx,
                   const double *ty, const double *tz, const double *cx, const double *cy,
                   const double *cz, int nsources, const double *sx, const double *sy,
                   const double *sz, const double *nx, const double *ny, const double *nz,
                   double scale_re, double scale_im, int accumulate, double *out) {
  SUMPY_REQUIRE(kernel_kind >= 0 && kernel_kind <= 5, "qbx_line: kernel kind %d", kernel_kind);
  const int dim = (kernel_kind == LAP3 || kernel_kind == HELM3 || kernel_kind == YUK3) ? 3 : 2;
  SUMPY_REQUIRE(order >= 0 && order <= QMAX, "qbx_line: order %d not in [0, %d]", order, QMAX);
  SUMPY_REQUIRE(src_kind >= 0 && src_kind <= 2, "qbx_line: source kind %d", src_kind);
  SUMPY_REQUIRE(src_kind != 2 || (src_axis >= 0 && src_axis < dim),
                "qbx_line: source axis %d", src_axis);
  SUMPY_REQUIRE(tgt_axis >= -1 && tgt_axis < dim, "qbx_line: target axis %d", tgt_axis);
  SUMPY_REQUIRE(ntargets >= 0 && nsources >= 0, "qbx_line: negative sizes");
  SUMPY_REQUIRE(out && tx && ty && cx && cy && (dim == 2 || (tz && cz)),
                "qbx_line: null target/centre/output pointer");
  SUMPY_REQUIRE(nsources == 0 || (sx && sy && (dim == 2 || sz)), "qbx_line: null sources");
  SUMPY_REQUIRE(src_kind != 1 || nsources == 0 || (nx && ny && (dim == 2 || nz)),
                "qbx_line: the directional source derivative needs the directions");
  if (kernel_kind == YUK2 || kernel_kind == YUK3)
    SUMPY_REQUIRE(k_im == 0.0 && k_re > 0.0,
                  "qbx_line: kernel kind %d needs a real positive parameter", kernel_kind);
  if (kernel_kind == HELM2)
    SUMPY_REQUIRE(k_im > 0.0 ? k_re >= 0.0 : (k_im == 0.0 && k_re > 0.0),
                  "qbx_line: 2D Helmholtz needs k > 0 or Re k >= 0 < Im k");
  const double *T[3] = {tx, ty, tz}, *Cc[3] = {cx, cy, cz}, *S[3] = {sx, sy, sz},
               *N[3] = {nx, ny, nz};
  g.p = order;
  g.src_kind = src_kind;
  g.src_axis = src_axis;
  g.tgt_axis = tgt_axis;
  g.kk = {k_re, k_im};
  g.scale = {scale_re, scale_im};
  g.nt = ntargets;
  g.ns = nsources;
  for (int d = 0; d < 3; ++d) {
    g.t[d] = T[d];
    g.c[d] = Cc[d];
    g.s[d] = S[d];
    g.n[d] = N[d];
  }
  g.accumulate = accumulate;
  g.out = out;
  return SUMPYHIP_OK;
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" int sumpyhip_qbx_line(int kernel_kind, int order, int src_kind, int src_axis,
                                 int tgt_axis, double k_re, double k_im, int ntargets,
                                 const double *tx, const double *ty, const double *tz,
                                 const double *cx, const double *cy, const double *cz,
                                 int nsources, const double *sx, const double *sy,
                                 const double *sz, const double *nx, const double *ny,
                                 const double *nz, const double *density, double scale_re,
                                 double scale_im, int accumulate, double *result,
                                 void *stream) {
  LineArgs g;
  int rc = fill_line_args(g, kernel_kind, order, src_kind, src_axis, tgt_axis, k_re, k_im,
                          ntargets, tx, ty, tz, cx, cy, cz, nsources, sx, sy, sz, nx, ny, nz,
                          scale_re, scale_im, accumulate, result);
  if (rc) return rc;
  SUMPY_REQUIRE(nsources == 0 || density, "qbx_line: null density");
  if (ntargets == 0) return SUMPYHIP_OK;
  g.density = density;
  return dispatch_line(kernel_kind, g, -1, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int sumpyhip_qbx_line_matrix(int kernel_kind, int order, int src_kind,
                                        int src_axis, int tgt_axis, double k_re, double k_im,
                                        int ntargets, const double *tx, const double *ty,
                                        const double *tz, const double *cx, const double *cy,
                                        const double *cz, int nsources, const double *sx,
                                        const double *sy, const double *sz, const double *nx,
                                        const double *ny, const double *nz, long long nentries,
                                        const int *tgt_idx, const int *src_idx,
                                        double scale_re, double scale_im, int accumulate,
                                        double *out, void *stream) {
  LineArgs g;
  int rc = fill_line_args(g, kernel_kind, order, src_kind, src_axis, tgt_axis, k_re, k_im,
                          ntargets, tx, ty, tz, cx, cy, cz, nsources, sx, sy, sz, nx, ny, nz,
                          scale_re, scale_im, accumulate, out);
  if (rc) return rc;
  SUMPY_REQUIRE(nentries >= 0, "qbx_line_matrix: negative entry count");
  SUMPY_REQUIRE((tgt_idx == nullptr) == (src_idx == nullptr),
                "qbx_line_matrix: give both index arrays or neither");
  SUMPY_REQUIRE(tgt_idx || nentries == (long long)ntargets*nsources,
                "qbx_line_matrix: %lld entries for a %d x %d matrix", nentries, ntargets,
                nsources);
  SUMPY_REQUIRE((nentries + TILE - 1)/TILE < (1LL << 31), "qbx_line_matrix: too many entries");
  g.density = nullptr;
  return dispatch_line(kernel_kind, g, nentries, tgt_idx, src_idx, (hipStream_t)stream);
}
