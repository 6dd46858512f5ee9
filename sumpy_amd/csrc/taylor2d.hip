// 2D volume-Taylor expansions (full and PDE-compressed) of the Laplace
// (G = log r) and Helmholtz (G = H0(k r), real k) kernels.
//
// Reference: VolumeTaylorMultipoleExpansion / VolumeTaylorLocalExpansion and
// their LinearPDEConforming forms (sumpy/expansion/multipole.py:70-419,
// sumpy/expansion/local.py:200-510) with the PDE projection of
// LinearPDEBasedExpansionTermsWrangler (sumpy/expansion/__init__.py:563-886),
// as called by the operators of sumpy/p2e.py, e2e.py and e2p.py.  With
// a = centre - source, b = target - centre, d = target centre - source
// centre and rho the rscale of the expansion:
//   P2M  M_m = sum_s q_s (a/rho)^m / m!                 (then P^T, compressed)
//   M2P  u(t) = scale sum_m M_m rho^|m| d^m G(b)
//   P2L  L_m = sum_s q_s rho^|m| d^m G(a)               (stored m)
//   L2P  u(t) = scale sum_m (P L)_m (b/rho)^m / m!
//   M2M  M'_m = sum_{n<=m} M_n (rho/rho')^|n| ((c - c')/rho')^(m-n)/(m-n)!  (then P'^T)
//   L2L  L'_m = (rho'/rho)^|m| sum_{n>=m} (P L)_n ((c' - c)/rho)^(n-m)/(n-m)!
//   M2L  L_t = (rho_t/rho_s)^|t| sum_s M_s D[t + s],  D[m] = rho_s^|m| d^m G(d)
// where multi-indices are (m_x, m_y), the sums run over the stored indices of
// the expansions (every index of the full Taylor set; m_x <= 1 for the
// compressed set), and P maps stored to full local coefficients by the PDE:
// u_{m_x, m_y} = -u_{m_x-2, m_y+2} - (k rho)^2 u_{m_x-2, m_y} (k = 0 for
// Laplace).  A multipole compresses by P^T; any full preimage of a stored
// multipole (the stored values, zero elsewhere) translates to the same
// compressed result, since translation preserves the PDE.
//
// Derivatives of G: with W = d_x + i d_y and Wbar = d_x - i d_y,
//   d_x^mx d_y^my = 2^-n i^-my (W + Wbar)^mx (W - Wbar)^my,  n = mx + my,
// W Wbar = Laplacian (= -k^2 on H0(k r), 0 on log r), and
//   Wbar^n log r = (-1)^(n-1) 2^(n-1) (n-1)! / z^n (n >= 1), W^n log r its
//   conjugate (z = x + i y);
//   W^n H0(k r) = (-k)^n H_n(k r) e^{i n th}, Wbar^n H0(k r) = (-k)^n H_n(k r) e^{-i n th}.
// Bessel functions from bessel.h.  Scaling constants (-1/(2 pi), i/4) are
// applied by the evaluations (M2P, L2P) only, as in the reference.
//
// Layout: coefficients indexed internally by tri(mx, my) = n(n+1)/2 + my
// (degree-major, the reference's full storage order in 2D); rows of the
// expansion arrays hold the stored coefficients (real for Laplace, complex
// (re, im) for Helmholtz).  One 128-thread block per box / class; sizes are
// small (order <= 12: 91 full coefficients), so the kernels are latency-bound.
#include "common.h"
#include "bessel.h"

namespace sumpyhip {
namespace {

constexpr int TPMAX = 12;                               // highest order
constexpr int TNF = (TPMAX + 1)*(TPMAX + 2)/2;          // full coefficients at TPMAX
constexpr int TND = (2*TPMAX + 1)*(2*TPMAX + 2)/2;      // derivatives up to 2 TPMAX
constexpr int TTH = 128;                                // threads per block
constexpr int TCH = 64;                                 // sources per LDS chunk (P2L)

struct T2E {
  int order, compressed, nstored;
  const int *smx, *smy;
};

__device__ __forceinline__ int tri(int mx, int my) {
  const int n = mx + my;
  return n*(n + 1)/2 + my;
}
__device__ __forceinline__ void untri(int e, int &mx, int &my) {
  int n = 0;
  while ((n + 1)*(n + 2)/2 <= e) ++n;
  my = e - n*(n + 1)/2;
  mx = n - my;
}
__device__ __forceinline__ int nfull(int p) { return (p + 1)*(p + 2)/2; }

__device__ __forceinline__ cplx cfma_s(cplx a, double s, cplx acc) {
  return {acc.re + a.re*s, acc.im + a.im*s};
}
__device__ __forceinline__ cplx cfma(cplx a, cplx b, cplx acc) {
  return {acc.re + a.re*b.re - a.im*b.im, acc.im + a.re*b.im + a.im*b.re};
}

template <bool CX>
__device__ __forceinline__ cplx ld(const double *p, size_t i) {
  return CX ? cplx{p[2*i], p[2*i + 1]} : cplx{p[i], 0.0};
}
template <bool CX>
__device__ __forceinline__ void stv(double *p, size_t i, cplx v) {
  if (CX) {
    p[2*i] = v.re;
    p[2*i + 1] = v.im;
  } else {
    p[i] = v.re;
  }
}
template <bool CX>
__device__ __forceinline__ void addv(double *p, size_t i, cplx v) {
  if (CX) {
    p[2*i] += v.re;
    p[2*i + 1] += v.im;
  } else {
    p[i] += v.re;
  }
}

// Wp[n] = W^n G, Wm[n] = Wbar^n G at the scaled vector (dx, dy) (= d/rho),
// n = 0..N, for KIND 0 (Laplace, lrho = log rho enters the n = 0 term) or
// KIND 1 (Helmholtz, kr = k rho; complex k: Bessel functions of complex
// argument)
template <int KIND>
__device__ void wseq(double dx, double dy, cplx kr, double lrho, int N, cplx *Wp, cplx *Wm) {
  const double r2 = dx*dx + dy*dy, r = sqrt(r2);
  if (KIND == 0) {
    Wp[0] = Wm[0] = {log(r) + lrho, 0.0};
    const cplx iz = {dx/r2, -dy/r2};
    cplx zp = iz;
    double f = 1.0;
    for (int n = 1; n <= N; ++n) {
      const cplx v = cscale(zp, (n & 1) ? f : -f);
      Wm[n] = v;
      Wp[n] = {v.re, -v.im};
      zp = cmul(zp, iz);
      f *= 2.0*n;
    }
  } else {
    // H_n = J_n + i Y_n into Wp first
    if (kr.im == 0.0) {
      double J[2*TPMAX + 2], Y[2*TPMAX + 2];
      bessel_jy(kr.re*r, N < 1 ? 1 : N, J, Y);
      for (int n = 0; n <= N; ++n) Wp[n] = {J[n], Y[n]};
    } else {
      cplx J[2*TPMAX + 2], Y[2*TPMAX + 2];
      bessel_jy_c(cscale(kr, r), N < 1 ? 1 : N, J, Y);
      for (int n = 0; n <= N; ++n) Wp[n] = {J[n].re - Y[n].im, J[n].im + Y[n].re};
    }
    const cplx e1 = {dx/r, dy/r}, mkr = {-kr.re, -kr.im};
    cplx e = {1.0, 0.0}, c = {1.0, 0.0};
    for (int n = 0; n <= N; ++n) {
      const cplx h = cmul(Wp[n], c);
      Wp[n] = cmul(h, e);
      Wm[n] = cmul(h, cplx{e.re, -e.im});
      e = cmul(e, e1);
      c = cmul(c, mkr);
    }
  }
}

// d_x^mx d_y^my G from the W sequences; mk2 = -(k rho)^2 (Helmholtz)
template <int KIND>
__device__ cplx deriv(int mx, int my, const cplx *Wp, const cplx *Wm, cplx mk2) {
  const int n = mx + my;
  cplx acc = {0.0, 0.0};
  double cx = 1.0;  // C(mx, i)
  for (int i = 0; i <= mx; ++i) {
    double cy = 1.0;  // C(my, j)
    for (int j = 0; j <= my; ++j) {
      const int a = i + j, b = n - a;
      const double c = ((my - j) & 1) ? -cx*cy : cx*cy;
      if (KIND == 1) {
        const int lo = a < b ? a : b;
        cplx f = {c, 0.0};
        for (int q = 0; q < lo; ++q) f = cmul(f, mk2);
        acc = cfma(a >= b ? Wp[a - b] : Wm[b - a], f, acc);
      } else if (a == 0 || b == 0) {
        acc = cfma_s(a >= b ? Wp[a] : Wm[b], c, acc);
      }
      cy = cy*(my - j)/(j + 1);
    }
    cx = cx*(mx - i)/(i + 1);
  }
  acc = cscale(acc, ldexp(1.0, -n));
  switch (my & 3) {
    case 0: return acc;
    case 1: return {acc.im, -acc.re};
    case 2: return {-acc.re, -acc.im};
    default: return {-acc.im, acc.re};
  }
}

// stored -> full local coefficients in F (stored positions already set)
__device__ void proj_fwd(int p, cplx kap, cplx *F) {
  for (int mx = 2; mx <= p; ++mx) {
    for (int my = threadIdx.x; my <= p - mx; my += blockDim.x) {
      const cplx a = F[tri(mx - 2, my + 2)], b = F[tri(mx - 2, my)];
      const cplx kb = cmul(kap, b);
      F[tri(mx, my)] = {-a.re - kb.re, -a.im - kb.im};
    }
    __syncthreads();
  }
}

// full multipole moments in F -> P^T F at the stored positions
__device__ void proj_bwd(int p, cplx kap, cplx *F) {
  for (int mx = p; mx >= 2; --mx) {
    for (int my = threadIdx.x; my <= p - mx; my += blockDim.x) {
      const cplx v = F[tri(mx, my)];
      cplx &o = F[tri(mx - 2, my + 2)];
      o = {o.re - v.re, o.im - v.im};
    }
    __syncthreads();
    for (int my = threadIdx.x; my <= p - mx; my += blockDim.x) {
      const cplx v = F[tri(mx, my)];
      cplx &o = F[tri(mx - 2, my)];
      const cplx kv = cmul(kap, v);
      o = {o.re - kv.re, o.im - kv.im};
    }
    __syncthreads();
  }
}

__device__ __forceinline__ double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}

// {{{ P2M

// Dipole weights u (wdx, wdy; NULL: none) contribute u . grad_d G(t - s): the
// derivative of the moments in a = c - s, (u_x/rho) (a/rho)^(m - e_x)/(m - e_x)!
// and the same in y.
template <bool CX>
__global__ __launch_bounds__(TTH) void t2_p2m(T2E ex, const int *source_boxes,
                                              const int *bss, const int *bsc,
                                              const double *centers, int ncenters,
                                              const double *sx, const double *sy,
                                              const double *w, const double *wdx,
                                              const double *wdy, double rscale, cplx kap,
                                              double *out, int base) {
  __shared__ cplx F[TNF];
  __shared__ double px[TCH][TPMAX + 1], py[TCH][TPMAX + 1];
  __shared__ cplx sw[TCH], swx[TCH], swy[TCH];
  const bool dip = wdx || wdy;
  const int box = source_boxes[blockIdx.x], p = ex.order, nf = nfull(p);
  const double cx = centers[box], cy = centers[ncenters + box];
  const int s0 = bss[box], ns = bsc[box];
  cplx acc = {0.0, 0.0};
  int mx = 0, my = 0;
  const int e = threadIdx.x;
  if (e < nf) untri(e, mx, my);
  for (int c0 = 0; c0 < ns; c0 += TCH) {
    __syncthreads();
    if (threadIdx.x < TCH && c0 + threadIdx.x < ns) {
      const int s = s0 + c0 + threadIdx.x;
      const double ax = (cx - sx[s])/rscale, ay = (cy - sy[s])/rscale;
      double fx = 1.0, fy = 1.0;
      for (int i = 0; i <= p; ++i) {
        px[threadIdx.x][i] = fx;
        py[threadIdx.x][i] = fy;
        fx *= ax/(i + 1);
        fy *= ay/(i + 1);
      }
      sw[threadIdx.x] = ld<CX>(w, s);
      if (dip) {
        swx[threadIdx.x] = wdx ? cscale(ld<CX>(wdx, s), 1.0/rscale) : cplx{0.0, 0.0};
        swy[threadIdx.x] = wdy ? cscale(ld<CX>(wdy, s), 1.0/rscale) : cplx{0.0, 0.0};
      }
    }
    __syncthreads();
    const int cnt = min(TCH, ns - c0);
    if (e < nf) {
      for (int j = 0; j < cnt; ++j) acc = cfma_s(sw[j], px[j][mx]*py[j][my], acc);
      if (dip)
        for (int j = 0; j < cnt; ++j) {
          if (mx > 0) acc = cfma_s(swx[j], px[j][mx - 1]*py[j][my], acc);
          if (my > 0) acc = cfma_s(swy[j], px[j][mx]*py[j][my - 1], acc);
        }
    }
  }
  for (int i = e; i < TNF; i += TTH) F[i] = {0.0, 0.0};
  __syncthreads();
  if (e < nf) F[e] = acc;
  __syncthreads();
  if (ex.compressed) proj_bwd(p, kap, F);
  double *row = out + (size_t)(box - base)*ex.nstored*(CX ? 2 : 1);
  for (int j = threadIdx.x; j < ex.nstored; j += TTH) stv<CX>(row, j, F[tri(ex.smx[j], ex.smy[j])]);
}

// }}}

// {{{ L2P, M2P

// out (potential), gx, gy (d/dx, d/dy: the local polynomial differentiated,
// (1/rho) F_{m + e} (b/rho)^m/m!); any may be NULL
template <bool CX>
__global__ __launch_bounds__(TTH) void t2_l2p(T2E ex, const int *target_boxes, const int *bts,
                                              const int *btc, const double *centers,
                                              int ncenters, const double *tx, const double *ty,
                                              double rscale, cplx kap, const double *src,
                                              int base, cplx scale, double *out, double *gx,
                                              double *gy) {
  __shared__ cplx F[TNF];
  const int box = target_boxes[blockIdx.x], p = ex.order;
  for (int i = threadIdx.x; i < TNF; i += TTH) F[i] = {0.0, 0.0};
  __syncthreads();
  const double *row = src + (size_t)(box - base)*ex.nstored*(CX ? 2 : 1);
  for (int j = threadIdx.x; j < ex.nstored; j += TTH) F[tri(ex.smx[j], ex.smy[j])] = ld<CX>(row, j);
  __syncthreads();
  if (ex.compressed) proj_fwd(p, kap, F);
  const double cx = centers[box], cy = centers[ncenters + box];
  const int t0 = bts[box], nt = btc[box];
  for (int it = threadIdx.x; it < nt; it += TTH) {
    const int t = t0 + it;
    const double bx = (tx[t] - cx)/rscale, by = (ty[t] - cy)/rscale;
    double fy[TPMAX + 1];
    fy[0] = 1.0;
    for (int i = 1; i <= p; ++i) fy[i] = fy[i - 1]*by/i;
    cplx acc = {0.0, 0.0}, ax = {0.0, 0.0}, ay = {0.0, 0.0};
    double fx = 1.0, fxm = 0.0;  // (b_x/rho)^mx/mx!, the same at mx - 1
    for (int mx = 0; mx <= p; ++mx) {
      cplx s = {0.0, 0.0}, sy = {0.0, 0.0};
      for (int my = 0; my <= p - mx; ++my) {
        s = cfma_s(F[tri(mx, my)], fy[my], s);
        if (my > 0) sy = cfma_s(F[tri(mx, my)], fy[my - 1], sy);
      }
      acc = cfma_s(s, fx, acc);
      ax = cfma_s(s, fxm, ax);
      ay = cfma_s(sy, fx, ay);
      fxm = fx;
      fx *= bx/(mx + 1);
    }
    if (out) stv<CX>(out, t, cmul(acc, scale));
    if (gx) stv<CX>(gx, t, cmul(cscale(ax, 1.0/rscale), scale));
    if (gy) stv<CX>(gy, t, cmul(cscale(ay, 1.0/rscale), scale));
  }
}

// out, gx, gy as in t2_l2p (added to): d/dx of rho^|m| d^m G is
// rho^|m| d^(m + e_x) G
template <int KIND>
__global__ __launch_bounds__(TTH) void t2_m2p(T2E ex, const int *target_boxes, const int *starts,
                                              const int *lists, const int *bts, const int *btc,
                                              const double *centers, int ncenters,
                                              const double *tx, const double *ty,
                                              double rscale, cplx k, const double *src,
                                              int base, cplx scale, double *out, double *gx,
                                              double *gy) {
  const bool grad = gx || gy;
  constexpr bool CX = KIND == 1;
  __shared__ cplx M[TNF];
  const int it = blockIdx.x, box = target_boxes[it];
  const int t0 = bts[box], nt = btc[box];
  const cplx kr = cscale(k, rscale), mk2 = cscale(cmul(kr, kr), -1.0);
  const double lrho = log(rscale);
  // targets of this thread (the box has at most a few hundred)
  for (int tb = 0; tb < nt; tb += TTH) {
    const int t = t0 + tb + threadIdx.x;
    const bool act = tb + threadIdx.x < nt;
    cplx acc = {0.0, 0.0}, ax = {0.0, 0.0}, ay = {0.0, 0.0};
    for (int q = starts[it]; q < starts[it + 1]; ++q) {
      const int sb = lists[q];
      __syncthreads();
      const double *row = src + (size_t)(sb - base)*ex.nstored*(CX ? 2 : 1);
      for (int j = threadIdx.x; j < ex.nstored; j += TTH) M[j] = ld<CX>(row, j);
      __syncthreads();
      if (!act) continue;
      cplx Wp[TPMAX + 2], Wm[TPMAX + 2];
      wseq<KIND>((tx[t] - centers[sb])/rscale, (ty[t] - centers[ncenters + sb])/rscale, kr,
                 lrho, ex.order + (grad ? 1 : 0), Wp, Wm);
      for (int j = 0; j < ex.nstored; ++j) {
        const int mx = ex.smx[j], my = ex.smy[j];
        if (out) acc = cfma(M[j], deriv<KIND>(mx, my, Wp, Wm, mk2), acc);
        if (grad) {
          ax = cfma(M[j], deriv<KIND>(mx + 1, my, Wp, Wm, mk2), ax);
          ay = cfma(M[j], deriv<KIND>(mx, my + 1, Wp, Wm, mk2), ay);
        }
      }
    }
    if (act) {
      if (out) addv<CX>(out, t, cmul(acc, scale));
      if (gx) addv<CX>(gx, t, cmul(cscale(ax, 1.0/rscale), scale));
      if (gy) addv<CX>(gy, t, cmul(cscale(ay, 1.0/rscale), scale));
    }
  }
}

// }}}

// {{{ P2L

// Dipole weights u (NULL: none): L_m += (u . grad) rho^|m| d^m G(c - s), i.e.
// (u_x rho^|m| d^(m + e_x) G + u_y ...), from the derivatives up to order p + 1
template <int KIND>
__global__ __launch_bounds__(TTH) void t2_p2l(T2E ex, const int *target_boxes, const int *starts,
                                              const int *lists, const int *bss, const int *bsc,
                                              const double *centers, int ncenters,
                                              const double *sx, const double *sy,
                                              const double *w, const double *wdx,
                                              const double *wdy, double rscale, cplx k,
                                              double *out, int base) {
  constexpr bool CX = KIND == 1;
  __shared__ cplx Wp[TCH][TPMAX + 2], Wm[TCH][TPMAX + 2];
  __shared__ cplx sw[TCH], swx[TCH], swy[TCH];
  const bool dip = wdx || wdy;
  const int it = blockIdx.x, box = target_boxes[it];
  const double cx = centers[box], cy = centers[ncenters + box];
  const cplx kr = cscale(k, rscale), mk2 = cscale(cmul(kr, kr), -1.0);
  const double lrho = log(rscale);
  const int j = threadIdx.x;
  const bool act = j < ex.nstored;
  const int mx = act ? ex.smx[j] : 0, my = act ? ex.smy[j] : 0;
  cplx acc = {0.0, 0.0};
  for (int q = starts[it]; q < starts[it + 1]; ++q) {
    const int sb = lists[q], s0 = bss[sb], ns = bsc[sb];
    for (int c0 = 0; c0 < ns; c0 += TCH) {
      __syncthreads();
      if (threadIdx.x < TCH && c0 + threadIdx.x < ns) {
        const int s = s0 + c0 + threadIdx.x;
        wseq<KIND>((cx - sx[s])/rscale, (cy - sy[s])/rscale, kr, lrho, ex.order + (dip ? 1 : 0),
                   Wp[threadIdx.x], Wm[threadIdx.x]);
        sw[threadIdx.x] = ld<CX>(w, s);
        if (dip) {
          swx[threadIdx.x] = wdx ? cscale(ld<CX>(wdx, s), 1.0/rscale) : cplx{0.0, 0.0};
          swy[threadIdx.x] = wdy ? cscale(ld<CX>(wdy, s), 1.0/rscale) : cplx{0.0, 0.0};
        }
      }
      __syncthreads();
      const int cnt = min(TCH, ns - c0);
      if (act)
        for (int s = 0; s < cnt; ++s) {
          acc = cfma(sw[s], deriv<KIND>(mx, my, Wp[s], Wm[s], mk2), acc);
          if (dip) {
            acc = cfma(swx[s], deriv<KIND>(mx + 1, my, Wp[s], Wm[s], mk2), acc);
            acc = cfma(swy[s], deriv<KIND>(mx, my + 1, Wp[s], Wm[s], mk2), acc);
          }
        }
    }
  }
  // (the reference's P2L keeps multi-index rows up to TTH >= 91 stored)
  if (act) stv<CX>(out + (size_t)(box - base)*ex.nstored*(CX ? 2 : 1), j, acc);
}

// }}}

// {{{ M2M, L2L

template <bool CX>
__global__ __launch_bounds__(TTH) void t2_m2m(T2E es, T2E et, const int *target_boxes,
                                              const int *child_ids, int child_stride,
                                              const double *centers, int ncenters,
                                              double src_rscale, double tgt_rscale,
                                              cplx kap_t, const double *src, int sbase,
                                              double *tgt, int tbase) {
  __shared__ cplx G[TNF], Ms[TNF];
  __shared__ double dxp[TPMAX + 1], dyp[TPMAX + 1], rr[TPMAX + 1];
  const int par = target_boxes[blockIdx.x], pt = et.order, nf = nfull(pt);
  const double cpx = centers[par], cpy = centers[ncenters + par];
  const int e = threadIdx.x;
  int mx = 0, my = 0;
  if (e < nf) untri(e, mx, my);
  cplx acc = {0.0, 0.0};
  for (int q = 0; q < 4; ++q) {
    const int ch = child_ids[(size_t)q*child_stride + par];
    if (ch == 0) continue;
    __syncthreads();
    const double *row = src + (size_t)(ch - sbase)*es.nstored*(CX ? 2 : 1);
    for (int j = threadIdx.x; j < es.nstored; j += TTH) Ms[j] = ld<CX>(row, j);
    if (threadIdx.x == 0) {
      const double Dx = (cpx - centers[ch])/tgt_rscale,
                   Dy = (cpy - centers[ncenters + ch])/tgt_rscale;
      const double ratio = src_rscale/tgt_rscale;
      double fx = 1.0, fy = 1.0, f = 1.0;
      for (int i = 0; i <= TPMAX; ++i) {
        dxp[i] = fx;
        dyp[i] = fy;
        rr[i] = f;
        fx *= Dx/(i + 1);
        fy *= Dy/(i + 1);
        f *= ratio;
      }
    }
    __syncthreads();
    if (e < nf)
      for (int j = 0; j < es.nstored; ++j) {
        const int nx = es.smx[j], ny = es.smy[j];
        if (nx <= mx && ny <= my)
          acc = cfma_s(Ms[j], rr[nx + ny]*dxp[mx - nx]*dyp[my - ny], acc);
      }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < TNF; i += TTH) G[i] = {0.0, 0.0};
  __syncthreads();
  if (e < nf) G[e] = acc;
  __syncthreads();
  if (et.compressed) proj_bwd(pt, kap_t, G);
  double *row = tgt + (size_t)(par - tbase)*et.nstored*(CX ? 2 : 1);
  for (int j = threadIdx.x; j < et.nstored; j += TTH) addv<CX>(row, j, G[tri(et.smx[j], et.smy[j])]);
}

template <bool CX>
__global__ __launch_bounds__(TTH) void t2_l2l(T2E es, T2E et, const int *target_boxes,
                                              const int *parent_ids, const double *centers,
                                              int ncenters, double src_rscale,
                                              double tgt_rscale, cplx kap_s,
                                              const double *src, int sbase, double *tgt,
                                              int tbase) {
  __shared__ cplx F[TNF];
  __shared__ double dxp[TPMAX + 1], dyp[TPMAX + 1];
  const int ch = target_boxes[blockIdx.x], par = parent_ids[ch], ps = es.order;
  for (int i = threadIdx.x; i < TNF; i += TTH) F[i] = {0.0, 0.0};
  __syncthreads();
  const double *row = src + (size_t)(par - sbase)*es.nstored*(CX ? 2 : 1);
  for (int j = threadIdx.x; j < es.nstored; j += TTH) F[tri(es.smx[j], es.smy[j])] = ld<CX>(row, j);
  if (threadIdx.x == 0) {
    const double Dx = (centers[ch] - centers[par])/src_rscale,
                 Dy = (centers[ncenters + ch] - centers[ncenters + par])/src_rscale;
    double fx = 1.0, fy = 1.0;
    for (int i = 0; i <= TPMAX; ++i) {
      dxp[i] = fx;
      dyp[i] = fy;
      fx *= Dx/(i + 1);
      fy *= Dy/(i + 1);
    }
  }
  __syncthreads();
  if (es.compressed) proj_fwd(ps, kap_s, F);
  const double ratio = tgt_rscale/src_rscale;
  double *out = tgt + (size_t)(ch - tbase)*et.nstored*(CX ? 2 : 1);
  for (int j = threadIdx.x; j < et.nstored; j += TTH) {
    const int mx = et.smx[j], my = et.smy[j];
    cplx acc = {0.0, 0.0};
    for (int nx = mx; nx <= ps; ++nx)
      for (int ny = my; nx + ny <= ps; ++ny)
        acc = cfma_s(F[tri(nx, ny)], dxp[nx - mx]*dyp[ny - my], acc);
    addv<CX>(out, j, cscale(acc, ipow(ratio, mx + my)));
  }
}

// }}}

// {{{ M2L

// D[tri(m)] = rho^|m| d^m G(d), |m| <= N, block-cooperative (ends with a barrier)
template <int KIND>
__device__ void deriv_table(double dx, double dy, double rscale, cplx k, int N, cplx *Wp,
                            cplx *Wm, cplx *D) {
  const cplx kr = cscale(k, rscale), mk2 = cscale(cmul(kr, kr), -1.0);
  if (threadIdx.x == 0) wseq<KIND>(dx/rscale, dy/rscale, kr, log(rscale), N, Wp, Wm);
  __syncthreads();
  for (int e = threadIdx.x; e < nfull(N); e += blockDim.x) {
    int mx, my;
    untri(e, mx, my);
    D[e] = deriv<KIND>(mx, my, Wp, Wm, mk2);
  }
  __syncthreads();
}

template <int KIND>
__global__ __launch_bounds__(TTH) void t2_m2l_data(int N, cplx k, int ncls, const double *vecs,
                                                   int nvecs, int class_start, double rscale,
                                                   double *out) {
  constexpr bool CX = KIND == 1;
  __shared__ cplx Wp[2*TPMAX + 1], Wm[2*TPMAX + 1], D[TND];
  const int c = blockIdx.x, v = class_start + c;
  deriv_table<KIND>(vecs[v], vecs[nvecs + v], rscale, k, N, Wp, Wm, D);
  double *row = out + (size_t)c*nfull(N)*(CX ? 2 : 1);
  for (int e = threadIdx.x; e < nfull(N); e += TTH) stv<CX>(row, e, D[e]);
}

template <int KIND>
__global__ __launch_bounds__(TTH) void t2_m2l(T2E es, T2E et, cplx k, const int *target_boxes,
                                              const int *starts, const int *lists,
                                              const int *classes, int class_start,
                                              const double *data, const double *centers,
                                              int ncenters, double src_rscale,
                                              double tgt_rscale, const double *src, int sbase,
                                              double *tgt, int tbase) {
  constexpr bool CX = KIND == 1;
  __shared__ cplx Wp[2*TPMAX + 1], Wm[2*TPMAX + 1], D[TND], Ms[TNF];
  const int it = blockIdx.x, box = target_boxes[it];
  const int N = es.order + et.order, nd = nfull(N);
  const int j = threadIdx.x;
  const bool act = j < et.nstored;
  const int tx = act ? et.smx[j] : 0, ty = act ? et.smy[j] : 0;
  cplx acc = {0.0, 0.0};
  for (int q = starts[it]; q < starts[it + 1]; ++q) {
    const int sb = lists[q];
    __syncthreads();
    const double *row = src + (size_t)(sb - sbase)*es.nstored*(CX ? 2 : 1);
    for (int i = threadIdx.x; i < es.nstored; i += TTH) Ms[i] = ld<CX>(row, i);
    if (data) {
      const double *crow = data + (size_t)(classes[q] - class_start)*nd*(CX ? 2 : 1);
      for (int i = threadIdx.x; i < nd; i += TTH) D[i] = ld<CX>(crow, i);
      __syncthreads();
    } else {
      deriv_table<KIND>(centers[box] - centers[sb],
                        centers[ncenters + box] - centers[ncenters + sb], src_rscale, k, N, Wp,
                        Wm, D);
    }
    if (act)
      for (int s = 0; s < es.nstored; ++s)
        acc = cfma(Ms[s], D[tri(tx + es.smx[s], ty + es.smy[s])], acc);
  }
  if (act)
    stv<CX>(tgt + (size_t)(box - tbase)*et.nstored*(CX ? 2 : 1), j,
            cscale(acc, ipow(tgt_rscale/src_rscale, tx + ty)));
}

// }}}

int check_expn(const sumpyhip_t2d_expn *e, const char *what) {
  SUMPY_REQUIRE(e, "%s: null expansion", what);
  SUMPY_REQUIRE(e->order >= 0 && e->order <= TPMAX, "%s: order %d not in [0, %d]", what,
                e->order, TPMAX);
  const int nf = (e->order + 1)*(e->order + 2)/2;
  SUMPY_REQUIRE(e->nstored >= 1 && e->nstored <= nf && e->stored_mx && e->stored_my,
                "%s: bad stored index table", what);
  SUMPY_REQUIRE(e->nstored <= TTH, "%s: %d stored coefficients > %d", what, e->nstored, TTH);
  return SUMPYHIP_OK;
}

// (k rho)^2 on the host
cplx ksq(cplx k, double rho) {
  const double a = k.re*rho, b = k.im*rho;
  return {a*a - b*b, 2.0*a*b};
}

T2E to_t2e(const sumpyhip_t2d_expn *e) {
  return T2E{e->order, e->compressed, e->nstored, e->stored_mx, e->stored_my};
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

#define T2_KIND(kind, what)                                                          \
  SUMPY_REQUIRE(kind == 0 || kind == 1, "%s: kind %d (0 Laplace, 1 Helmholtz)", what, kind); \
  SUMPY_REQUIRE(kind == 0 ? k_im == 0.0 : (k_im >= 0.0 && (k > 0.0 || k_im > 0.0)),     \
                "%s: wavenumber %g%+gi (Helmholtz: Im k >= 0, k != 0; Laplace: real)", what, k, \
                k_im);                                                                   \
  const cplx kk = {kind ? k : 0.0, kind ? k_im : 0.0}

extern "C" int sumpyhip_t2d_p2m(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int nboxes,
                                const int *source_boxes, const int *box_source_starts,
                                const int *box_source_counts, const double *centers,
                                int ncenters, const double *sx, const double *sy,
                                const double *weights, const double *weights_dx,
                                const double *weights_dy, double rscale,
                                double *tgt_expansions, int tgt_base_ibox, void *stream) {
  T2_KIND(kind, "t2d_p2m");
  int rc = check_expn(ex, "t2d_p2m");
  if (rc) return rc;
  if (nboxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(nboxes > 0 && source_boxes && box_source_starts && box_source_counts && centers &&
                    sx && sy && weights && tgt_expansions && rscale > 0,
                "t2d_p2m: bad arguments");
  const cplx kap = ksq(kk, rscale);
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_p2m<true><<<nboxes, TTH, 0, st>>>(to_t2e(ex), source_boxes, box_source_starts,
                                         box_source_counts, centers, ncenters, sx, sy, weights,
                                         weights_dx, weights_dy, rscale, kap, tgt_expansions,
                                         tgt_base_ibox);
  else
    t2_p2m<false><<<nboxes, TTH, 0, st>>>(to_t2e(ex), source_boxes, box_source_starts,
                                          box_source_counts, centers, ncenters, sx, sy, weights,
                                          weights_dx, weights_dy, rscale, kap, tgt_expansions,
                                          tgt_base_ibox);
  return check_launch("t2d_p2m");
}

extern "C" int sumpyhip_t2d_l2p(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *box_target_starts,
                                const int *box_target_counts, const double *centers,
                                int ncenters, const double *tx, const double *ty, double rscale,
                                const double *src_expansions, int src_base_ibox,
                                double scale_re, double scale_im, double *result,
                                double *result_dx, double *result_dy, void *stream) {
  T2_KIND(kind, "t2d_l2p");
  int rc = check_expn(ex, "t2d_l2p");
  if (rc) return rc;
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt > 0 && target_boxes && box_target_starts && box_target_counts && centers &&
                    tx && ty && src_expansions && (result || result_dx || result_dy) && rscale > 0,
                "t2d_l2p: bad arguments");
  const cplx kap = ksq(kk, rscale);
  const cplx sc = {scale_re, scale_im};
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_l2p<true><<<ntgt, TTH, 0, st>>>(to_t2e(ex), target_boxes, box_target_starts,
                                       box_target_counts, centers, ncenters, tx, ty, rscale, kap,
                                       src_expansions, src_base_ibox, sc, result, result_dx,
                                       result_dy);
  else
    t2_l2p<false><<<ntgt, TTH, 0, st>>>(to_t2e(ex), target_boxes, box_target_starts,
                                        box_target_counts, centers, ncenters, tx, ty, rscale,
                                        kap, src_expansions, src_base_ibox, sc, result, result_dx,
                                        result_dy);
  return check_launch("t2d_l2p");
}

extern "C" int sumpyhip_t2d_m2p(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *source_box_starts,
                                const int *source_box_lists, const int *box_target_starts,
                                const int *box_target_counts, const double *centers,
                                int ncenters, const double *tx, const double *ty, double rscale,
                                const double *src_expansions, int src_base_ibox,
                                double scale_re, double scale_im, double *result,
                                double *result_dx, double *result_dy, void *stream) {
  T2_KIND(kind, "t2d_m2p");
  int rc = check_expn(ex, "t2d_m2p");
  if (rc) return rc;
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt > 0 && target_boxes && source_box_starts && source_box_lists &&
                    box_target_starts && box_target_counts && centers && tx && ty &&
                    src_expansions && (result || result_dx || result_dy) && rscale > 0,
                "t2d_m2p: bad arguments");
  const cplx sc = {scale_re, scale_im};
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_m2p<1><<<ntgt, TTH, 0, st>>>(to_t2e(ex), target_boxes, source_box_starts,
                                    source_box_lists, box_target_starts, box_target_counts,
                                    centers, ncenters, tx, ty, rscale, kk, src_expansions,
                                    src_base_ibox, sc, result, result_dx, result_dy);
  else
    t2_m2p<0><<<ntgt, TTH, 0, st>>>(to_t2e(ex), target_boxes, source_box_starts,
                                    source_box_lists, box_target_starts, box_target_counts,
                                    centers, ncenters, tx, ty, rscale, kk, src_expansions,
                                    src_base_ibox, sc, result, result_dx, result_dy);
  return check_launch("t2d_m2p");
}

extern "C" int sumpyhip_t2d_p2l(int kind, const sumpyhip_t2d_expn *ex, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *source_box_starts,
                                const int *source_box_lists, const int *box_source_starts,
                                const int *box_source_counts, const double *centers,
                                int ncenters, const double *sx, const double *sy,
                                const double *weights, const double *weights_dx,
                                const double *weights_dy, double rscale,
                                double *tgt_expansions, int tgt_base_ibox, void *stream) {
  T2_KIND(kind, "t2d_p2l");
  int rc = check_expn(ex, "t2d_p2l");
  if (rc) return rc;
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt > 0 && target_boxes && source_box_starts && source_box_lists &&
                    box_source_starts && box_source_counts && centers && sx && sy && weights &&
                    tgt_expansions && rscale > 0,
                "t2d_p2l: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_p2l<1><<<ntgt, TTH, 0, st>>>(to_t2e(ex), target_boxes, source_box_starts,
                                    source_box_lists, box_source_starts, box_source_counts,
                                    centers, ncenters, sx, sy, weights, weights_dx, weights_dy,
                                    rscale, kk,
                                    tgt_expansions, tgt_base_ibox);
  else
    t2_p2l<0><<<ntgt, TTH, 0, st>>>(to_t2e(ex), target_boxes, source_box_starts,
                                    source_box_lists, box_source_starts, box_source_counts,
                                    centers, ncenters, sx, sy, weights, weights_dx, weights_dy,
                                    rscale, kk,
                                    tgt_expansions, tgt_base_ibox);
  return check_launch("t2d_p2l");
}

extern "C" int sumpyhip_t2d_m2m(int kind, const sumpyhip_t2d_expn *src_ex,
                                const sumpyhip_t2d_expn *tgt_ex, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *box_child_ids,
                                int child_stride, const double *centers, int ncenters,
                                double src_rscale, double tgt_rscale,
                                const double *src_expansions, int src_base_ibox,
                                double *tgt_expansions, int tgt_base_ibox, void *stream) {
  T2_KIND(kind, "t2d_m2m");
  int rc = check_expn(src_ex, "t2d_m2m");
  if (!rc) rc = check_expn(tgt_ex, "t2d_m2m");
  if (rc) return rc;
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt > 0 && target_boxes && box_child_ids && centers && src_expansions &&
                    tgt_expansions && src_rscale > 0 && tgt_rscale > 0,
                "t2d_m2m: bad arguments");
  const cplx kap = ksq(kk, tgt_rscale);
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_m2m<true><<<ntgt, TTH, 0, st>>>(to_t2e(src_ex), to_t2e(tgt_ex), target_boxes,
                                       box_child_ids, child_stride, centers, ncenters,
                                       src_rscale, tgt_rscale, kap, src_expansions,
                                       src_base_ibox, tgt_expansions, tgt_base_ibox);
  else
    t2_m2m<false><<<ntgt, TTH, 0, st>>>(to_t2e(src_ex), to_t2e(tgt_ex), target_boxes,
                                        box_child_ids, child_stride, centers, ncenters,
                                        src_rscale, tgt_rscale, kap, src_expansions,
                                        src_base_ibox, tgt_expansions, tgt_base_ibox);
  return check_launch("t2d_m2m");
}

extern "C" int sumpyhip_t2d_l2l(int kind, const sumpyhip_t2d_expn *src_ex,
                                const sumpyhip_t2d_expn *tgt_ex, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *box_parent_ids,
                                const double *centers, int ncenters, double src_rscale,
                                double tgt_rscale, const double *src_expansions,
                                int src_base_ibox, double *tgt_expansions, int tgt_base_ibox,
                                void *stream) {
  T2_KIND(kind, "t2d_l2l");
  int rc = check_expn(src_ex, "t2d_l2l");
  if (!rc) rc = check_expn(tgt_ex, "t2d_l2l");
  if (rc) return rc;
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt > 0 && target_boxes && box_parent_ids && centers && src_expansions &&
                    tgt_expansions && src_rscale > 0 && tgt_rscale > 0,
                "t2d_l2l: bad arguments");
  const cplx kap = ksq(kk, src_rscale);
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_l2l<true><<<ntgt, TTH, 0, st>>>(to_t2e(src_ex), to_t2e(tgt_ex), target_boxes,
                                       box_parent_ids, centers, ncenters, src_rscale,
                                       tgt_rscale, kap, src_expansions, src_base_ibox,
                                       tgt_expansions, tgt_base_ibox);
  else
    t2_l2l<false><<<ntgt, TTH, 0, st>>>(to_t2e(src_ex), to_t2e(tgt_ex), target_boxes,
                                        box_parent_ids, centers, ncenters, src_rscale,
                                        tgt_rscale, kap, src_expansions, src_base_ibox,
                                        tgt_expansions, tgt_base_ibox);
  return check_launch("t2d_l2l");
}

extern "C" int sumpyhip_t2d_m2l_data(int kind, int order_sum, double k, double k_im, int ncls,
                                     const double *vecs, int nvecs, int class_start,
                                     double rscale, double *out, void *stream) {
  T2_KIND(kind, "t2d_m2l_data");
  SUMPY_REQUIRE(order_sum >= 0 && order_sum <= 2*TPMAX, "t2d_m2l_data: order sum %d",
                order_sum);
  if (ncls == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ncls > 0 && vecs && out && rscale > 0 && class_start >= 0 &&
                    class_start + ncls <= nvecs,
                "t2d_m2l_data: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_m2l_data<1><<<ncls, TTH, 0, st>>>(order_sum, kk, ncls, vecs, nvecs, class_start, rscale,
                                         out);
  else
    t2_m2l_data<0><<<ncls, TTH, 0, st>>>(order_sum, kk, ncls, vecs, nvecs, class_start, rscale,
                                         out);
  return check_launch("t2d_m2l_data");
}

extern "C" int sumpyhip_t2d_m2l(int kind, const sumpyhip_t2d_expn *src_ex,
                                const sumpyhip_t2d_expn *tgt_ex, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *src_box_starts,
                                const int *src_box_lists, const int *classes, int class_start,
                                const double *class_data, const double *centers, int ncenters,
                                double src_rscale, double tgt_rscale,
                                const double *src_expansions, int src_base_ibox,
                                double *tgt_expansions, int tgt_base_ibox, void *stream) {
  T2_KIND(kind, "t2d_m2l");
  int rc = check_expn(src_ex, "t2d_m2l");
  if (!rc) rc = check_expn(tgt_ex, "t2d_m2l");
  if (rc) return rc;
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt > 0 && target_boxes && src_box_starts && src_box_lists && centers &&
                    src_expansions && tgt_expansions && src_rscale > 0 && tgt_rscale > 0,
                "t2d_m2l: bad arguments");
  SUMPY_REQUIRE(!class_data || classes, "t2d_m2l: class data without class lists");
  hipStream_t st = (hipStream_t)stream;
  if (kind)
    t2_m2l<1><<<ntgt, TTH, 0, st>>>(to_t2e(src_ex), to_t2e(tgt_ex), kk, target_boxes,
                                    src_box_starts, src_box_lists, classes, class_start,
                                    class_data, centers, ncenters, src_rscale, tgt_rscale,
                                    src_expansions, src_base_ibox, tgt_expansions,
                                    tgt_base_ibox);
  else
    t2_m2l<0><<<ntgt, TTH, 0, st>>>(to_t2e(src_ex), to_t2e(tgt_ex), kk, target_boxes,
                                    src_box_starts, src_box_lists, classes, class_start,
                                    class_data, centers, ncenters, src_rscale, tgt_rscale,
                                    src_expansions, src_base_ibox, tgt_expansions,
                                    tgt_base_ibox);
  return check_launch("t2d_m2l");
}
