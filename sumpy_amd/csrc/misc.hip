// Error reporting, device queries, source-weight and target-output
// combination kernels, and the 2D QBX layer potential.
#include "common.h"

#include <cmath>
#include <cstring>
#include <mutex>

namespace sumpyhip {

static thread_local char g_err[1024];

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int num_cus() {
  static int cached = 0;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!cached) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cached = prop.multiProcessorCount;
    else
      cached = 256;
  }
  return cached;
}

namespace {

// {{{ weights and outputs

__global__ void axpy_weights_kernel(int is_complex, int n, double coeff, const double *q,
                                    const double *dir, double *w) {
  for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < n; i += gridDim.x*blockDim.x) {
    const double f = coeff*(dir ? dir[i] : 1.0);
    if (is_complex) {
      w[2*i] += f*q[2*i];
      w[2*i + 1] += f*q[2*i + 1];
    } else {
      w[i] += f*q[i];
    }
  }
}

constexpr int MAXTERMS = 8;
struct CombineArgs {
  int is_complex, n, nterms, accumulate;
  const double *ins[MAXTERMS];
  double coeffs[MAXTERMS];
  const double *mults[MAXTERMS];
  double *out;
};

__global__ void combine_kernel(CombineArgs a) {
  for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < a.n; i += gridDim.x*blockDim.x) {
    double re = 0, im = 0;
    for (int j = 0; j < a.nterms; ++j) {
      const double f = a.coeffs[j]*(a.mults[j] ? a.mults[j][i] : 1.0);
      if (a.is_complex) {
        re += f*a.ins[j][2*i];
        im += f*a.ins[j][2*i + 1];
      } else {
        re += f*a.ins[j][i];
      }
    }
    if (a.is_complex) {
      if (a.accumulate) { a.out[2*i] += re; a.out[2*i + 1] += im; }
      else { a.out[2*i] = re; a.out[2*i + 1] = im; }
    } else {
      if (a.accumulate) a.out[i] += re;
      else a.out[i] = re;
    }
  }
}

// }}}

// {{{ QBX: 2D Laplace line-Taylor, complex-analytic closed forms

constexpr int QB = 256;

// One (target, source) term of the line-Taylor sum: a = c - s, b = t - c,
// z = b/a; source kind 0 = SLP, 1 = DLP with normal n, 2/3 = x/y source
// derivative; target axis -1 = potential, 0/1 = x/y target derivative.
__device__ __forceinline__ double qbx2d_pair(int order, int src_kind, int tgt_axis,
                                             double are, double aim, double bre, double bim,
                                             double nx, double ny) {
  const double aa = are*are + aim*aim;
  const double ia_re = are/aa, ia_im = -aim/aa;  // 1/a
  const double zre = bre*ia_re - bim*ia_im, zim = bre*ia_im + bim*ia_re;
  if (src_kind == 0 && tgt_axis < 0) {
    // log a + sum_{i=1}^p (-1)^{i-1} z^i / i   (real part)
    double hre = 0, him = 0;
    for (int i = order; i >= 1; --i) {
      const double c = ((i & 1) ? 1.0 : -1.0)/i;
      const double nre = hre*zre - him*zim + c, nim = hre*zim + him*zre;
      hre = nre; him = nim;
    }
    return 0.5*log(aa) + (hre*zre - him*zim);  // times z
  }
  // value = Re[ w * sum_{i=0}^p c_i (-z)^i ]
  //   SLP + d/da: w = 1/a, c_i = 1 (times e_axis factor)
  //   DLP/axis source deriv: w = -n/a, c_i = 1;  with d/da: w = n/a^2, c_i = i+1
  double wre, wim;
  int lin = 0;
  if (src_kind == 0) {
    wre = ia_re; wim = ia_im;
  } else {
    double nre, nim;
    if (src_kind == 1) { nre = nx; nim = ny; }
    else if (src_kind == 2) { nre = 1; nim = 0; }
    else { nre = 0; nim = 1; }
    if (tgt_axis < 0) {
      wre = -(nre*ia_re - nim*ia_im); wim = -(nre*ia_im + nim*ia_re);
    } else {
      const double a2re = ia_re*ia_re - ia_im*ia_im, a2im = 2*ia_re*ia_im;
      wre = nre*a2re - nim*a2im; wim = nre*a2im + nim*a2re;
      lin = 1;
    }
  }
  double hre = 0, him = 0;
  const double mzre = -zre, mzim = -zim;
  for (int i = order; i >= 0; --i) {
    const double c = lin ? double(i + 1) : 1.0;
    const double nre = hre*mzre - him*mzim + c, nim = hre*mzim + him*mzre;
    hre = nre; him = nim;
  }
  double pre = wre*hre - wim*him, pim = wre*him + wim*hre;
  if (tgt_axis == 1) pre = -pim;  // times i
  return pre;
}

// {{{ forward and adjoint kernels: order, layer kind and target axis at
// compile time, sources split over lanes

// 1/x for x > 0 without a division: hardware reciprocal refined by two
// Newton steps (relative error at the FP64 rounding level)
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// The pair term of qbx2d_pair with the order P, the layer kind (DLP: the
// source has a direction n, for the double layer or an axis derivative) and
// the target axis TA fixed; the Horner loops are unrolled.
template <int P, bool DLP, int TA>
__device__ __forceinline__ double qbx_pair_t(double are, double aim, double bre, double bim,
                                             double nre, double nim) {
  const double aa = are*are + aim*aim;
  const double inv = rcp64(aa);
  const double ia_re = are*inv, ia_im = -aim*inv;  // 1/a
  const double zre = bre*ia_re - bim*ia_im, zim = bre*ia_im + bim*ia_re;
  if constexpr (!DLP && TA < 0) {
    double hre = 0, him = 0;
    #pragma unroll
    for (int i = P; i >= 1; --i) {
      const double c = ((i & 1) ? 1.0 : -1.0)/i;
      const double nre2 = hre*zre - him*zim + c, nim2 = hre*zim + him*zre;
      hre = nre2; him = nim2;
    }
    return 0.5*log(aa) + (hre*zre - him*zim);
  } else {
    double wre, wim;
    constexpr bool lin = DLP && TA >= 0;
    if constexpr (!DLP) {
      wre = ia_re; wim = ia_im;
    } else if constexpr (TA < 0) {
      wre = -(nre*ia_re - nim*ia_im); wim = -(nre*ia_im + nim*ia_re);
    } else {
      const double a2re = ia_re*ia_re - ia_im*ia_im, a2im = 2*ia_re*ia_im;
      wre = nre*a2re - nim*a2im; wim = nre*a2im + nim*a2re;
    }
    double hre = 0, him = 0;
    const double mzre = -zre, mzim = -zim;
    #pragma unroll
    for (int i = P; i >= 0; --i) {
      const double c = lin ? double(i + 1) : 1.0;
      const double nre2 = hre*mzre - him*mzim + c, nim2 = hre*mzim + him*mzre;
      hre = nre2; him = nim2;
    }
    if constexpr (TA == 1) return -(wre*him + wim*hre);
    return wre*hre - wim*him;
  }
}

// P >= 0: compile-time order; P < 0: the run-time order (any order)
template <int P, bool DLP, int TA>
__device__ __forceinline__ double qbx_pair_any(int order, double are, double aim, double bre,
                                               double bim, double nre, double nim) {
  if constexpr (P >= 0) return qbx_pair_t<P, DLP, TA>(are, aim, bre, bim, nre, nim);
  else return qbx2d_pair(order, DLP ? 1 : 0, TA, are, aim, bre, bim, nre, nim);
}

constexpr int QT = 16;   // targets (or sources, transposed) per block
constexpr int QL = 16;   // lanes sharing one target; they split the sources

// sum over the QL lanes of a group (butterfly in a fixed order: the sum does
// not depend on scheduling)
__device__ __forceinline__ double group_sum(double v) {
  #pragma unroll
  for (int m = QL/2; m >= 1; m >>= 1) v += __shfl_xor(v, m, QL);
  return v;
}

// out[t] = scale sum_s dens_s term(t, s).  Block: QT targets x QL lanes; the
// block's sources are staged QB at a time in LDS and lane j of a target takes
// sources j, j + QL, ...  Source kind: DLP with normals (nx, ny), or with the
// constant direction (dirx, diry) when nx is null (axis source derivatives).
template <int P, bool DLP, int TA>
__global__ __launch_bounds__(QB) void qbx2d_fwd(int order, int nt, const double *tx,
                                                const double *ty,
                                                const double *cx, const double *cy, int ns,
                                                const double *sx, const double *sy,
                                                const double *nx, const double *ny,
                                                double dirx, double diry, const double *dens,
                                                double scale, double *out) {
  __shared__ double s_x[QB], s_y[QB], s_nx[QB], s_ny[QB], s_d[QB];
  const int tid = threadIdx.x, g = tid/QL, j = tid - g*QL;
  const int tr = blockIdx.x*QT + g;
  const int t = min(tr, nt - 1);
  const double ccx = cx[t], ccy = cy[t];
  const double bre = tx[t] - ccx, bim = ty[t] - ccy;
  double res = 0;
  for (int sc = 0; sc < ns; sc += QB) {
    __syncthreads();
    if (sc + tid < ns) {
      const int s = sc + tid;
      s_x[tid] = sx[s]; s_y[tid] = sy[s]; s_d[tid] = dens[s];
      if (DLP) { s_nx[tid] = nx ? nx[s] : dirx; s_ny[tid] = nx ? ny[s] : diry; }
    }
    __syncthreads();
    const int nn = min(QB, ns - sc);
    if (nn == QB) {
      #pragma unroll 2
      for (int m = j; m < QB; m += QL)
        res += s_d[m]*qbx_pair_any<P, DLP, TA>(order, ccx - s_x[m], ccy - s_y[m], bre, bim,
                                                DLP ? s_nx[m] : 0.0, DLP ? s_ny[m] : 0.0);
    } else {
      for (int m = j; m < nn; m += QL)
        res += s_d[m]*qbx_pair_any<P, DLP, TA>(order, ccx - s_x[m], ccy - s_y[m], bre, bim,
                                                DLP ? s_nx[m] : 0.0, DLP ? s_ny[m] : 0.0);
    }
  }
  res = group_sum(res);
  if (j == 0 && tr < nt) out[t] = scale*res;
}

// }}}

// Matrix entries (reference LayerPotentialMatrixGenerator and
// LayerPotentialMatrixSubsetGenerator, sumpy/qbx.py:342-539): entry e is the
// pair (tidx[e], sidx[e]), or (e / ns, e % ns) for the full row-major matrix;
// out[e] (+)= scale * term (accumulate: one call per source kernel).
__global__ __launch_bounds__(QB) void qbx2d_matrix_kernel(
    int order, int src_kind, int tgt_axis, int ns, long long nentries, const int *tidx,
    const int *sidx, const double *tx, const double *ty, const double *cx, const double *cy,
    const double *sx, const double *sy, const double *nx, const double *ny, double scale,
    int accumulate, double *out) {
  for (long long e = blockIdx.x*(long long)QB + threadIdx.x; e < nentries;
       e += (long long)gridDim.x*QB) {
    int t, s;
    if (tidx) { t = tidx[e]; s = sidx[e]; }
    else { t = (int)(e/ns); s = (int)(e - (long long)t*ns); }
    const double val = qbx2d_pair(order, src_kind, tgt_axis, cx[t] - sx[s], cy[t] - sy[s],
                                  tx[t] - cx[t], ty[t] - cy[t],
                                  src_kind == 1 ? nx[s] : 0.0, src_kind == 1 ? ny[s] : 0.0);
    out[e] = (accumulate ? out[e] : 0.0) + scale*val;
  }
}

// }}}

// {{{ QBX adjoints (gradients of sum_t g_t out_t)

// Line-Taylor sums of the 2D Laplace SLP / DLP and their a- and b-derivatives
// (complex notation, a = c - s, b = t - c, u = -b/a, p = order):
//   SLP: F = log a + sum_{i=1}^p (-1)^{i-1} (b/a)^i / i,
//        F_a = (1/a) sum_{i<=p} u^i,        F_b = (1/a) sum_{i<p} u^i
//   DLP: F = n G,  G = -(1/a) sum_{i<=p} u^i,
//        G_a = (1/a^2) sum_{i<=p} (i+1) u^i,  G_b = (1/a^2) sum_{i<p} (i+1) u^i
// The kernel value is Re F; for a real 2-vector x entering through a complex
// variable w, d Re F / dx = (Re F_w, -Im F_w).
struct QbxTerms {
  double fre;            // Re F
  double fa_re, fa_im;   // F_a
  double fb_re, fb_im;   // F_b
  double g_re, g_im;     // DLP: G (= dF/dn)
};

__device__ __forceinline__ QbxTerms qbx_terms(int order, bool dlp, double are, double aim,
                                              double bre, double bim, double nre, double nim) {
  const double aa = are*are + aim*aim;
  const double ia_re = are/aa, ia_im = -aim/aa;  // 1/a
  const double zre = bre*ia_re - bim*ia_im, zim = bre*ia_im + bim*ia_re;
  const double ure = -zre, uim = -zim;
  // Horner sums of sum_{i<=p} c_i u^i and sum_{i<p} c_i u^i (c_i = 1 or i + 1)
  double s0re = 0, s0im = 0, s1re = 0, s1im = 0;  // i <= p and i < p
  for (int i = order; i >= 0; --i) {
    const double c = dlp ? double(i + 1) : 1.0;
    double tre = s0re*ure - s0im*uim + c, tim = s0re*uim + s0im*ure;
    s0re = tre; s0im = tim;
    if (i < order) {
      tre = s1re*ure - s1im*uim + c; tim = s1re*uim + s1im*ure;
      s1re = tre; s1im = tim;
    }
  }
  QbxTerms r;
  if (!dlp) {
    r.fa_re = ia_re*s0re - ia_im*s0im; r.fa_im = ia_re*s0im + ia_im*s0re;
    r.fb_re = ia_re*s1re - ia_im*s1im; r.fb_im = ia_re*s1im + ia_im*s1re;
    // Re F = log|a| + Re sum_{i=1}^p (-1)^{i-1} z^i / i
    double hre = 0, him = 0;
    for (int i = order; i >= 1; --i) {
      const double c = ((i & 1) ? 1.0 : -1.0)/i;
      const double tre = hre*zre - him*zim + c, tim = hre*zim + him*zre;
      hre = tre; him = tim;
    }
    r.fre = 0.5*log(aa) + (hre*zre - him*zim);
    r.g_re = r.g_im = 0;
  } else {
    // sum_{i<=p} u^i for G itself
    double hre = 0, him = 0;
    for (int i = order; i >= 0; --i) {
      const double tre = hre*ure - him*uim + 1.0, tim = hre*uim + him*ure;
      hre = tre; him = tim;
    }
    r.g_re = -(ia_re*hre - ia_im*him); r.g_im = -(ia_re*him + ia_im*hre);
    r.fre = nre*r.g_re - nim*r.g_im;
    const double a2re = ia_re*ia_re - ia_im*ia_im, a2im = 2*ia_re*ia_im;  // 1/a^2
    const double gare = a2re*s0re - a2im*s0im, gaim = a2re*s0im + a2im*s0re;
    const double gbre = a2re*s1re - a2im*s1im, gbim = a2re*s1im + a2im*s1re;
    r.fa_re = nre*gare - nim*gaim; r.fa_im = nre*gaim + nim*gare;
    r.fb_re = nre*gbre - nim*gbim; r.fb_im = nre*gbim + nim*gbre;
  }
  return r;
}

// line-Taylor sums and their a- and b-derivatives with the order fixed
// (qbx_terms below is the run-time-order form)
template <int P, bool DLP>
__device__ __forceinline__ void qbx_terms_t(int order, double are, double aim, double bre,
                                            double bim, double nre, double nim, double &fre,
                                            double &fa_re, double &fa_im, double &fb_re,
                                            double &fb_im, double &g_re, double &g_im) {
  if constexpr (P < 0) {
    const QbxTerms r = qbx_terms(order, DLP, are, aim, bre, bim, nre, nim);
    fre = r.fre; fa_re = r.fa_re; fa_im = r.fa_im; fb_re = r.fb_re; fb_im = r.fb_im;
    g_re = r.g_re; g_im = r.g_im;
    return;
  }
  constexpr int PP = P < 0 ? 0 : P;
  const double aa = are*are + aim*aim;
  const double inv = rcp64(aa);
  const double ia_re = are*inv, ia_im = -aim*inv;
  const double zre = bre*ia_re - bim*ia_im, zim = bre*ia_im + bim*ia_re;
  const double ure = -zre, uim = -zim;
  // s1 = sum_{i<p} c_i u^i, s0 = s1 + c_p u^p (c_i = 1, or i + 1 for the DLP)
  double s1re = 0, s1im = 0;
  #pragma unroll
  for (int i = PP - 1; i >= 0; --i) {
    const double c = DLP ? double(i + 1) : 1.0;
    const double tre = s1re*ure - s1im*uim + c, tim = s1re*uim + s1im*ure;
    s1re = tre; s1im = tim;
  }
  double upre = 1, upim = 0;  // u^p
  #pragma unroll
  for (int i = 0; i < PP; ++i) {
    const double tre = upre*ure - upim*uim, tim = upre*uim + upim*ure;
    upre = tre; upim = tim;
  }
  const double cp = DLP ? double(PP + 1) : 1.0;
  const double s0re = s1re + cp*upre, s0im = s1im + cp*upim;
  if constexpr (!DLP) {
    fa_re = ia_re*s0re - ia_im*s0im; fa_im = ia_re*s0im + ia_im*s0re;
    fb_re = ia_re*s1re - ia_im*s1im; fb_im = ia_re*s1im + ia_im*s1re;
    double hre = 0, him = 0;
    #pragma unroll
    for (int i = PP; i >= 1; --i) {
      const double c = ((i & 1) ? 1.0 : -1.0)/i;
      const double tre = hre*zre - him*zim + c, tim = hre*zim + him*zre;
      hre = tre; him = tim;
    }
    fre = 0.5*log(aa) + (hre*zre - him*zim);
    g_re = g_im = 0;
  } else {
    // sum_{i<=p} u^i
    double hre = 0, him = 0;
    #pragma unroll
    for (int i = PP; i >= 0; --i) {
      const double tre = hre*ure - him*uim + 1.0, tim = hre*uim + him*ure;
      hre = tre; him = tim;
    }
    g_re = -(ia_re*hre - ia_im*him); g_im = -(ia_re*him + ia_im*hre);
    fre = nre*g_re - nim*g_im;
    const double a2re = ia_re*ia_re - ia_im*ia_im, a2im = 2*ia_re*ia_im;
    const double gare = a2re*s0re - a2im*s0im, gaim = a2re*s0im + a2im*s0re;
    const double gbre = a2re*s1re - a2im*s1im, gbim = a2re*s1im + a2im*s1re;
    fa_re = nre*gare - nim*gaim; fa_im = nre*gaim + nim*gare;
    fb_re = nre*gbre - nim*gbim; fb_im = nre*gbim + nim*gbre;
  }
}

// per target: d/dt and d/dc of g_t out_t (same block shape as qbx2d_fwd)
template <int P, bool DLP>
__global__ __launch_bounds__(QB) void qbx2d_grad_targets_t(
    int order, int nt, const double *tx, const double *ty, const double *cx, const double *cy, int ns,
    const double *sx, const double *sy, const double *nx, const double *ny, const double *dens,
    const double *g, double scale, double *dtx, double *dty, double *dcx, double *dcy) {
  __shared__ double s_x[QB], s_y[QB], s_nx[QB], s_ny[QB], s_d[QB];
  const int tid = threadIdx.x, gi = tid/QL, j = tid - gi*QL;
  const int tr = blockIdx.x*QT + gi;
  const int t = min(tr, nt - 1);
  const double ccx = cx[t], ccy = cy[t];
  const double bre = tx[t] - ccx, bim = ty[t] - ccy;
  double ab_re = 0, ab_im = 0, aa_re = 0, aa_im = 0;
  for (int sc = 0; sc < ns; sc += QB) {
    __syncthreads();
    if (sc + tid < ns) {
      const int s = sc + tid;
      s_x[tid] = sx[s]; s_y[tid] = sy[s]; s_d[tid] = dens[s];
      if (DLP) { s_nx[tid] = nx[s]; s_ny[tid] = ny[s]; }
    }
    __syncthreads();
    const int nn = min(QB, ns - sc);
    for (int m = j; m < nn; m += QL) {
      double fre, fare, faim, fbre, fbim, gre, gim;
      qbx_terms_t<P, DLP>(order, ccx - s_x[m], ccy - s_y[m], bre, bim, DLP ? s_nx[m] : 0.0,
                          DLP ? s_ny[m] : 0.0, fre, fare, faim, fbre, fbim, gre, gim);
      ab_re += s_d[m]*fbre; ab_im += s_d[m]*fbim;
      aa_re += s_d[m]*fare; aa_im += s_d[m]*faim;
    }
  }
  ab_re = group_sum(ab_re); ab_im = group_sum(ab_im);
  aa_re = group_sum(aa_re); aa_im = group_sum(aa_im);
  if (j == 0 && tr < nt) {
    const double w = scale*g[t];
    if (dtx) { dtx[t] = w*ab_re; dty[t] = -w*ab_im; }
    if (dcx) { dcx[t] = w*(aa_re - ab_re); dcy[t] = -w*(aa_im - ab_im); }
  }
}

// per source: d/dsigma, d/ds and (DLP) d/dn of sum_t g_t out_t; QT sources per
// block, the lanes of a source split the targets
template <int P, bool DLP>
__global__ __launch_bounds__(QB) void qbx2d_grad_sources_t(
    int order, int ns, const double *sx, const double *sy, const double *nx, const double *ny,
    const double *dens, int nt, const double *tx, const double *ty, const double *cx,
    const double *cy, const double *g, double scale, double *ddens, double *dsx, double *dsy,
    double *dnx, double *dny) {
  __shared__ double s_tx[QB], s_ty[QB], s_cx[QB], s_cy[QB], s_g[QB];
  const int tid = threadIdx.x, gi = tid/QL, j = tid - gi*QL;
  const int sr = blockIdx.x*QT + gi;
  const int s = min(sr, ns - 1);
  const double px = sx[s], py = sy[s];
  const double nre = DLP ? nx[s] : 0.0, nim = DLP ? ny[s] : 0.0;
  double val = 0, fa_re = 0, fa_im = 0, g_re = 0, g_im = 0;
  for (int tc = 0; tc < nt; tc += QB) {
    __syncthreads();
    if (tc + tid < nt) {
      const int t = tc + tid;
      s_tx[tid] = tx[t]; s_ty[tid] = ty[t]; s_cx[tid] = cx[t]; s_cy[tid] = cy[t];
      s_g[tid] = g[t];
    }
    __syncthreads();
    const int nn = min(QB, nt - tc);
    for (int m = j; m < nn; m += QL) {
      double fre, fare, faim, fbre, fbim, gre, gim;
      qbx_terms_t<P, DLP>(order, s_cx[m] - px, s_cy[m] - py, s_tx[m] - s_cx[m], s_ty[m] - s_cy[m],
                          nre, nim, fre, fare, faim, fbre, fbim, gre, gim);
      const double w = s_g[m];
      val += w*fre;
      fa_re += w*fare; fa_im += w*faim;
      g_re += w*gre; g_im += w*gim;
    }
  }
  val = group_sum(val);
  fa_re = group_sum(fa_re); fa_im = group_sum(fa_im);
  g_re = group_sum(g_re); g_im = group_sum(g_im);
  if (j == 0 && sr < ns) {
    const double sig = scale*dens[s];
    if (ddens) ddens[s] = scale*val;
    if (dsx) { dsx[s] = -sig*fa_re; dsy[s] = sig*fa_im; }
    if (DLP && dnx) { dnx[s] = sig*g_re; dny[s] = -sig*g_im; }
  }
}

// {{{ dispatch on (order, layer kind, target axis)

constexpr int QBX_PMAX = 10;  // orders 0..QBX_PMAX at compile time, others at run time

template <bool DLP, int TA, int P = 0>
void launch_qbx_fwd(int order, int grid, hipStream_t st, int nt, const double *tx,
                    const double *ty, const double *cx, const double *cy, int ns,
                    const double *sx, const double *sy, const double *nx, const double *ny,
                    double dirx, double diry, const double *dens, double scale, double *out) {
  if constexpr (P > QBX_PMAX) {
    qbx2d_fwd<-1, DLP, TA><<<grid, QB, 0, st>>>(order, nt, tx, ty, cx, cy, ns, sx, sy, nx, ny,
                                                 dirx, diry, dens, scale, out);
  } else {
    if (order == P)
      qbx2d_fwd<P, DLP, TA><<<grid, QB, 0, st>>>(order, nt, tx, ty, cx, cy, ns, sx, sy, nx, ny,
                                                  dirx, diry, dens, scale, out);
    else
      launch_qbx_fwd<DLP, TA, P + 1>(order, grid, st, nt, tx, ty, cx, cy, ns, sx, sy, nx, ny,
                                     dirx, diry, dens, scale, out);
  }
}

template <bool DLP, int P = 0>
void launch_qbx_grad_targets(int order, int grid, hipStream_t st, int nt, const double *tx,
                             const double *ty, const double *cx, const double *cy, int ns,
                             const double *sx, const double *sy, const double *nx,
                             const double *ny, const double *dens, const double *g,
                             double scale, double *dtx, double *dty, double *dcx, double *dcy) {
  if constexpr (P > QBX_PMAX) {
    qbx2d_grad_targets_t<-1, DLP><<<grid, QB, 0, st>>>(order, nt, tx, ty, cx, cy, ns, sx, sy,
                                                        nx, ny, dens, g, scale, dtx, dty, dcx,
                                                        dcy);
  } else {
    if (order == P)
      qbx2d_grad_targets_t<P, DLP><<<grid, QB, 0, st>>>(order, nt, tx, ty, cx, cy, ns, sx, sy,
                                                         nx, ny, dens, g, scale, dtx, dty, dcx,
                                                         dcy);
    else
      launch_qbx_grad_targets<DLP, P + 1>(order, grid, st, nt, tx, ty, cx, cy, ns, sx, sy, nx,
                                          ny, dens, g, scale, dtx, dty, dcx, dcy);
  }
}

template <bool DLP, int P = 0>
void launch_qbx_grad_sources(int order, int grid, hipStream_t st, int ns, const double *sx,
                             const double *sy, const double *nx, const double *ny,
                             const double *dens, int nt, const double *tx, const double *ty,
                             const double *cx, const double *cy, const double *g, double scale,
                             double *ddens, double *dsx, double *dsy, double *dnx, double *dny) {
  if constexpr (P > QBX_PMAX) {
    qbx2d_grad_sources_t<-1, DLP><<<grid, QB, 0, st>>>(order, ns, sx, sy, nx, ny, dens, nt, tx,
                                                        ty, cx, cy, g, scale, ddens, dsx, dsy,
                                                        dnx, dny);
  } else {
    if (order == P)
      qbx2d_grad_sources_t<P, DLP><<<grid, QB, 0, st>>>(order, ns, sx, sy, nx, ny, dens, nt, tx,
                                                         ty, cx, cy, g, scale, ddens, dsx, dsy,
                                                         dnx, dny);
    else
      launch_qbx_grad_sources<DLP, P + 1>(order, grid, st, ns, sx, sy, nx, ny, dens, nt, tx,
                                          ty, cx, cy, g, scale, ddens, dsx, dsy, dnx, dny);
  }
}

// }}}

// }}}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" const char *sumpyhip_last_error(void) { return g_err; }

extern "C" int sumpyhip_device_info(int *n_cu, int *major, int *minor) {
  int dev = 0;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return set_error(SUMPYHIP_EHIP, "device query: %s", hipGetErrorString(e));
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (major) *major = prop.major;
  if (minor) *minor = prop.minor;
  return SUMPYHIP_OK;
}

extern "C" int sumpyhip_axpy_weights(int is_complex, int n, double coeff,
                                     const double *strength, const double *dir, double *w,
                                     void *stream) {
  if (n == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(strength && w, "axpy_weights: null pointer");
  const int grid = (n + 255)/256 < 4096 ? (n + 255)/256 : 4096;
  axpy_weights_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(is_complex, n, coeff, strength,
                                                               dir, w);
  return check_launch("axpy_weights");
}

extern "C" int sumpyhip_combine_outputs(int is_complex, int n, int nterms,
                                        const double *const *ins, const double *coeffs,
                                        const double *const *mults, double *out,
                                        int accumulate, void *stream) {
  SUMPY_REQUIRE(nterms >= 1 && nterms <= MAXTERMS, "combine_outputs: 1..%d terms", MAXTERMS);
  if (n == 0) return SUMPYHIP_OK;
  CombineArgs a{};
  a.is_complex = is_complex; a.n = n; a.nterms = nterms; a.accumulate = accumulate;
  a.out = out;
  for (int j = 0; j < nterms; ++j) {
    a.ins[j] = ins[j];
    a.coeffs[j] = coeffs[j];
    a.mults[j] = mults ? mults[j] : nullptr;
  }
  const int grid = (n + 255)/256 < 4096 ? (n + 255)/256 : 4096;
  combine_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(a);
  return check_launch("combine_outputs");
}

extern "C" int sumpyhip_qbx2d(int order, int src_kind, int tgt_axis, int ntargets,
                              const double *tx, const double *ty, const double *cx,
                              const double *cy, int nsources, const double *sx,
                              const double *sy, const double *nx, const double *ny,
                              const double *density, double scale, double *out,
                              void *stream) {
  SUMPY_REQUIRE(order >= 0 && order <= 64, "qbx2d: order out of range");
  SUMPY_REQUIRE(src_kind >= 0 && src_kind <= 3, "qbx2d: unknown source kind %d", src_kind);
  SUMPY_REQUIRE(tgt_axis >= -1 && tgt_axis <= 1, "qbx2d: target axis must be -1, 0 or 1");
  if (src_kind == 1) SUMPY_REQUIRE(nx && ny, "qbx2d: DLP needs source normals");
  if (ntargets == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(nsources >= 1, "qbx2d: nsources must be >= 1");
  const int grid = (ntargets + QT - 1)/QT;
  hipStream_t st = (hipStream_t)stream;
  // source kinds 2 / 3 (x / y source derivative) are the double-layer form
  // with a constant direction
  const double dirx = src_kind == 2 ? 1.0 : 0.0, diry = src_kind == 3 ? 1.0 : 0.0;
  const double *nxp = src_kind == 1 ? nx : nullptr, *nyp = src_kind == 1 ? ny : nullptr;
  if (src_kind == 0) {
    if (tgt_axis < 0)
      launch_qbx_fwd<false, -1>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx, sy,
                                nullptr, nullptr, 0.0, 0.0, density, scale, out);
    else if (tgt_axis == 0)
      launch_qbx_fwd<false, 0>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx, sy,
                               nullptr, nullptr, 0.0, 0.0, density, scale, out);
    else
      launch_qbx_fwd<false, 1>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx, sy,
                               nullptr, nullptr, 0.0, 0.0, density, scale, out);
  } else {
    if (tgt_axis < 0)
      launch_qbx_fwd<true, -1>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx, sy,
                               nxp, nyp, dirx, diry, density, scale, out);
    else if (tgt_axis == 0)
      launch_qbx_fwd<true, 0>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx, sy,
                              nxp, nyp, dirx, diry, density, scale, out);
    else
      launch_qbx_fwd<true, 1>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx, sy,
                              nxp, nyp, dirx, diry, density, scale, out);
  }
  return check_launch("qbx2d");
}

extern "C" int sumpyhip_qbx2d_adjoint(int order, int src_kind, int ntargets, const double *tx,
                                      const double *ty, const double *cx, const double *cy,
                                      int nsources, const double *sx, const double *sy,
                                      const double *nx, const double *ny,
                                      const double *density, double scale,
                                      const double *grad_out, double *d_density, double *d_sx,
                                      double *d_sy, double *d_nx, double *d_ny, double *d_tx,
                                      double *d_ty, double *d_cx, double *d_cy, void *stream) {
  SUMPY_REQUIRE(order >= 0 && order <= 64, "qbx2d_adjoint: order out of range");
  SUMPY_REQUIRE(src_kind == 0 || src_kind == 1,
                "qbx2d_adjoint: single (0) or double (1) layer only, got %d", src_kind);
  if (src_kind == 1) SUMPY_REQUIRE(nx && ny, "qbx2d_adjoint: DLP needs source normals");
  SUMPY_REQUIRE(!d_sx == !d_sy && !d_nx == !d_ny && !d_tx == !d_ty && !d_cx == !d_cy,
                "qbx2d_adjoint: gradient outputs come in (x, y) pairs");
  if (ntargets == 0 || nsources == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(grad_out && density, "qbx2d_adjoint: null input");
  hipStream_t st = (hipStream_t)stream;
  if (d_tx || d_cx) {
    const int grid = (ntargets + QT - 1)/QT;
    if (src_kind == 1)
      launch_qbx_grad_targets<true>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx,
                                    sy, nx, ny, density, grad_out, scale, d_tx, d_ty, d_cx,
                                    d_cy);
    else
      launch_qbx_grad_targets<false>(order, grid, st, ntargets, tx, ty, cx, cy, nsources, sx,
                                     sy, nx, ny, density, grad_out, scale, d_tx, d_ty, d_cx,
                                     d_cy);
    int rc = check_launch("qbx2d_grad_targets");
    if (rc) return rc;
  }
  if (d_density || d_sx || d_nx) {
    const int grid = (nsources + QT - 1)/QT;
    if (src_kind == 1)
      launch_qbx_grad_sources<true>(order, grid, st, nsources, sx, sy, nx, ny, density,
                                    ntargets, tx, ty, cx, cy, grad_out, scale, d_density, d_sx,
                                    d_sy, d_nx, d_ny);
    else
      launch_qbx_grad_sources<false>(order, grid, st, nsources, sx, sy, nx, ny, density,
                                     ntargets, tx, ty, cx, cy, grad_out, scale, d_density, d_sx,
                                     d_sy, d_nx, d_ny);
    return check_launch("qbx2d_grad_sources");
  }
  return SUMPYHIP_OK;
}

extern "C" int sumpyhip_qbx2d_matrix(int order, int src_kind, int tgt_axis, int ntargets,
                                     const double *tx, const double *ty, const double *cx,
                                     const double *cy, int nsources, const double *sx,
                                     const double *sy, const double *nx, const double *ny,
                                     long long nentries, const int *tgtindices,
                                     const int *srcindices, double scale, int accumulate,
                                     double *out, void *stream) {
  SUMPY_REQUIRE(order >= 0 && order <= 64, "qbx2d_matrix: order out of range");
  SUMPY_REQUIRE(src_kind >= 0 && src_kind <= 3, "qbx2d_matrix: unknown source kind %d",
                src_kind);
  SUMPY_REQUIRE(tgt_axis >= -1 && tgt_axis <= 1, "qbx2d_matrix: target axis must be -1, 0 or 1");
  if (src_kind == 1) SUMPY_REQUIRE(nx && ny, "qbx2d_matrix: DLP needs source normals");
  SUMPY_REQUIRE((tgtindices == nullptr) == (srcindices == nullptr),
                "qbx2d_matrix: give both index arrays or neither");
  SUMPY_REQUIRE(tgtindices || nentries == (long long)ntargets*nsources,
                "qbx2d_matrix: the full matrix has ntargets * nsources entries");
  if (nentries == 0) return SUMPYHIP_OK;
  const long long nb = (nentries + QB - 1)/QB;
  const int grid = (int)(nb < 65536 ? nb : 65536);
  qbx2d_matrix_kernel<<<grid, QB, 0, (hipStream_t)stream>>>(
      order, src_kind, tgt_axis, nsources, nentries, tgtindices, srcindices, tx, ty, cx, cy,
      sx, sy, nx, ny, scale, accumulate, out);
  return check_launch("qbx2d_matrix");
}
