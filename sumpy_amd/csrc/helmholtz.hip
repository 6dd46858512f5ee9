// The kernel-dependent parts of the Helmholtz-3D FMM: translation-class data,
// M2P (list 3) and P2L (list 4) for G = e^{ik|d|}/|d| (reference
// sumpy/kernel.py:754-844), PDE-compressed Taylor expansions.
//
// Scaled derivatives.  With d = rho xi and kappa = k rho,
//   rho^|m| d^m G(d) = (1/rho) d_xi^m H(xi),   H(xi) = e^{i kappa |xi|}/|xi|.
// H is radial: H = F(s), s = |xi|^2/2, with F^(q) = g_q = ((1/r) d/dr)^q H, and
//   g_0 = e^{i kappa r}/r,  g_1 = e^{i kappa r}(i kappa r - 1)/r^3,
//   g_q = (-(2q - 1) g_{q-1} - kappa^2 g_{q-2})/r^2
// (the radial recurrence of sumpy/derivative_taker.py:354-382, after Tausch).
// Since s separates over the axes, the chain rule gives
//   d^(a,b,c) H = sum_{i,j,l} h(a,i) h(b,j) h(c,l) x^{2i-a} y^{2j-b} z^{2l-c} g_{i+j+l}
// with h(n,i) = n!/((2i-n)! (n-i)! 2^{n-i}), ceil(n/2) <= i <= n (the
// coefficients of d^n/dx^n F(x^2/2) in the F^(i)).  This is the sum the
// reference's RadialDerivativeTaker recursion (derivative_taker.py:290-346)
// expands to; here it is evaluated directly from a table of h.
//
// Complex values are interleaved (re, im) doubles, torch's complex128 layout.
// The kernel-independent stages (P2M, M2M, L2L, L2P) are the real kernels of
// expansions.hip and octant.hip applied to the real and imaginary planes with
// projection tables built for (k rho)^2; the M2L contraction is the real MFMA
// contraction of m2l.hip on the real embedding [Lr; Li] = [Ar -Ai; Ai Ar][Mr; Mi].
#include "common.h"

namespace sumpyhip {
namespace {

constexpr int HELM_MAXN = 52;  // derivative orders handled (class data: 2 p <= 48)
constexpr int HN = HELM_MAXN + 1;

// h(n, i) for n, i <= nmax into tab[n*(nmax + 1) + i] (zero outside
// ceil(n/2)..n); block-cooperative, ends with a barrier
__device__ void herm_table(double *tab, int nmax) {
  const int hs = nmax + 1;
  for (int n = threadIdx.x; n <= nmax; n += blockDim.x) {
    for (int i = 0; i <= nmax; ++i) tab[n*hs + i] = 0;
    double h = 1;  // h(n, ceil(n/2)): the product of the odd numbers <= n
    for (int o = 3; o <= n; o += 2) h *= o;
    for (int i = (n + 1)/2; i <= n; ++i) {
      tab[n*hs + i] = h;
      h *= 2.0*(n - i)/double((2*i + 2 - n)*(2*i + 1 - n));
    }
  }
  __syncthreads();
}

// g_0..g_n of e^{i kappa r}/r at r into (gre, gim)[q*stride]; kappa =
// kappa + i kappa_im (damped waves, kappa_im > 0: the same recurrence in
// complex arithmetic)
__device__ void helm_radial_c(double kappa, double kappa_im, double r, int n, double *gre,
                              double *gim, int stride) {
  const double ri = 1.0/r, ri2 = ri*ri;
  double sn, cs;
  sincos(kappa*r, &sn, &cs);
  const double damp = exp(-kappa_im*r);
  double ar = damp*cs*ri, ai = damp*sn*ri;
  gre[0] = ar;
  gim[0] = ai;
  if (n < 1) return;
  // e^{i kappa r}(i kappa r - 1)/r^3, i kappa r - 1 = (-kappa_im r - 1) + i kappa r
  const double ur = -kappa_im*r - 1.0, ui = kappa*r;
  double br = (ar*ur - ai*ui)*ri2, bi = (ar*ui + ai*ur)*ri2;
  gre[stride] = br;
  gim[stride] = bi;
  const double k2r = kappa*kappa - kappa_im*kappa_im, k2i = 2.0*kappa*kappa_im;
  for (int q = 2; q <= n; ++q) {
    const double f = -(2*q - 1);
    const double cr = (f*br - (k2r*ar - k2i*ai))*ri2, ci = (f*bi - (k2r*ai + k2i*ar))*ri2;
    gre[q*stride] = cr;
    gim[q*stride] = ci;
    ar = br; ai = bi; br = cr; bi = ci;
  }
}

__device__ void helm_radial(double kappa, double kappa_im, double r, int n, double *gre,
                            double *gim, int stride) {
  if (kappa_im != 0.0) {
    helm_radial_c(kappa, kappa_im, r, n, gre, gim, stride);
    return;
  }
  const double ri = 1.0/r, ri2 = ri*ri;
  double sn, cs;
  sincos(kappa*r, &sn, &cs);
  double ar = cs*ri, ai = sn*ri;
  gre[0] = ar;
  gim[0] = ai;
  if (n < 1) return;
  const double kr = kappa*r;
  // e^{i kr}(i kr - 1)/r^3
  double br = (-cs - sn*kr)*ri2*ri, bi = (cs*kr - sn)*ri2*ri;
  gre[stride] = br;
  gim[stride] = bi;
  const double k2 = kappa*kappa;
  for (int q = 2; q <= n; ++q) {
    const double f = -(2*q - 1);
    const double cr = (f*br - k2*ar)*ri2, ci = (f*bi - k2*ai)*ri2;
    gre[q*stride] = cr;
    gim[q*stride] = ci;
    ar = br; ai = bi; br = cr; bi = ci;
  }
}

// d^(a,b,c) H from the powers xp[e*ps] = x^e (yp, zp likewise), g_q at
// (gre, gim)[q*gs] and the h table of row stride hs
__device__ __forceinline__ cplx helm_deriv(int a, int b, int c, const double *xp,
                                           const double *yp, const double *zp, int ps,
                                           const double *gre, const double *gim, int gs,
                                           const double *ht, int hs) {
  double rr = 0, ri = 0;
  for (int i = (a + 1)/2; i <= a; ++i) {
    double sr = 0, si = 0;
    for (int j = (b + 1)/2; j <= b; ++j) {
      double tr = 0, ti = 0;
      for (int l = (c + 1)/2; l <= c; ++l) {
        const double f = ht[c*hs + l]*zp[(2*l - c)*ps];
        const int q = (i + j + l)*gs;
        tr += f*gre[q];
        ti += f*gim[q];
      }
      const double f = ht[b*hs + j]*yp[(2*j - b)*ps];
      sr += f*tr;
      si += f*ti;
    }
    const double f = ht[a*hs + i]*xp[(2*i - a)*ps];
    rr += f*sr;
    ri += f*si;
  }
  return {rr, ri};
}

// {{{ translation-class data

// out[c][0][off(m)] + i out[c][1][off(m)] = rho^|m| d^m G(v_c) for |m| <= N in
// the row-major box (M0, M1, M2); zero elsewhere
__global__ __launch_bounds__(256) void helm_class_data_kernel(
    int N, int M0, int M1, int M2, int nclasses, const double *vecs, int nvectors,
    int class_start, double rscale, double k, double kim, double *out) {
  __shared__ double ht[HN*HN];
  __shared__ double sp[3][HN], sg[2][HN];
  herm_table(ht, N);
  const size_t ncirc = (size_t)M0*M1*M2;
  const double rinv = 1.0/rscale;
  for (int c = blockIdx.x; c < nclasses; c += gridDim.x) {
    __syncthreads();  // the previous class is done with sp / sg
    if (threadIdx.x < 4) {
      const size_t v = class_start + c;
      const double x = vecs[v]*rinv, y = vecs[nvectors + v]*rinv,
                   z = vecs[2*(size_t)nvectors + v]*rinv;
      if (threadIdx.x < 3) {
        const double xi = threadIdx.x == 0 ? x : (threadIdx.x == 1 ? y : z);
        double pw = 1;
        for (int e = 0; e <= N; ++e) {
          sp[threadIdx.x][e] = pw;
          pw *= xi;
        }
      } else {
        helm_radial(k*rscale, kim*rscale, sqrt(x*x + y*y + z*z), N, sg[0], sg[1], 1);
      }
    }
    __syncthreads();
    double *o = out + (size_t)c*2*ncirc;
    for (size_t e = threadIdx.x; e < ncirc; e += blockDim.x) {
      const int mz = int(e % M2), my = int((e/M2) % M1), mx = int(e/((size_t)M2*M1));
      double re = 0, im = 0;
      if (mx + my + mz <= N) {
        const cplx d = helm_deriv(mx, my, mz, sp[0], sp[1], sp[2], 1, sg[0], sg[1], 1, ht, N + 1);
        re = d.re*rinv;
        im = d.im*rinv;
      }
      o[e] = re;
      o[ncirc + e] = im;
    }
  }
}

// }}}

// {{{ M2P

// One wave per target box; lane = target.  Per list-3 source box the wave
// stages the box's multipole (complex, stored order) in LDS, each lane fills
// its powers of xi = (t - c)/rho and g_q in its own LDS column, and
//   pot += scale sum_s M_s rho^|s| d^s G(t - c)
//   grad_i += scale sum_s M_s rho^|s| d^{s + e_i} G(t - c) / rho
// (reference sumpy/e2p.py:233-343, accumulates).
template <bool GRAD>
__global__ __launch_bounds__(64) void m2p_helm3d(
    sumpyhip_expn e, double k, double kim, int ntgt, const int *target_boxes, const int *starts,
    const int *lists, const int *bts, const int *btc, const double *centers, int nboxes,
    const double *tx, const double *ty, const double *tz, double rscale, double scale,
    const double *src, int sbase, double *pot, double *gx, double *gy, double *gz) {
  extern __shared__ double sh[];
  const int p = e.order, nst = e.nstored;
  const int NQ = p + 1 + (GRAD ? 1 : 0);  // highest derivative order
  const int lane = threadIdx.x;
  double *ht = sh;                      // [(NQ+1)^2]
  double *sM = ht + (NQ + 1)*(NQ + 1);  // [2*nst]
  double *tab = sM + 2*nst;             // [5*(NQ+1)][64]: x, y, z powers, g re, g im
  int *sid = reinterpret_cast<int *>(tab + 5*(NQ + 1)*64);  // [3*nst]
  herm_table(ht, NQ);
  for (int q = lane; q < 3*nst; q += 64) sid[q] = e.stored_ids[q];
  double *xp = tab + lane, *yp = xp + (NQ + 1)*64, *zp = yp + (NQ + 1)*64;
  double *gre = zp + (NQ + 1)*64, *gim = gre + (NQ + 1)*64;
  const double rinv = 1.0/rscale, kappa = k*rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    const int t0 = bts[tb], nt = btc[tb];
    for (int tc = 0; tc < nt; tc += 64) {
      const bool valid = tc + lane < nt;
      const int t = t0 + tc + (valid ? lane : 0);
      const double X = tx[t], Y = ty[t], Z = tz[t];
      double pr = 0, pi = 0, ar[3] = {0, 0, 0}, ai[3] = {0, 0, 0};
      for (int q = starts[ib]; q < starts[ib + 1]; ++q) {
        const int sb = lists[q];
        __syncthreads();
        const double *M = src + 2*(size_t)(sb - sbase)*nst;
        for (int u = lane; u < 2*nst; u += 64) sM[u] = M[u];
        const double x = (X - centers[sb])*rinv, y = (Y - centers[nboxes + sb])*rinv,
                     z = (Z - centers[2*(size_t)nboxes + sb])*rinv;
        double px = 1, py = 1, pz = 1;
        for (int u = 0; u <= NQ; ++u) {
          xp[u*64] = px; yp[u*64] = py; zp[u*64] = pz;
          px *= x; py *= y; pz *= z;
        }
        helm_radial(kappa, kim*rscale, sqrt(x*x + y*y + z*z), NQ, gre, gim, 64);
        __syncthreads();
        for (int s = 0; s < nst; ++s) {
          const int a = sid[s], b = sid[nst + s], c = sid[2*nst + s];
          const double mr = sM[2*s], mi = sM[2*s + 1];
          const cplx d = helm_deriv(a, b, c, xp, yp, zp, 64, gre, gim, 64, ht, NQ + 1);
          pr += mr*d.re - mi*d.im;
          pi += mr*d.im + mi*d.re;
          if (GRAD) {
            #pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
              const cplx g = helm_deriv(a + (ax == 0), b + (ax == 1), c + (ax == 2), xp, yp,
                                        zp, 64, gre, gim, 64, ht, NQ + 1);
              ar[ax] += mr*g.re - mi*g.im;
              ai[ax] += mr*g.im + mi*g.re;
            }
          }
        }
      }
      if (valid) {
        const double f = scale*rinv;
        if (pot) {
          pot[2*(size_t)t] += f*pr;
          pot[2*(size_t)t + 1] += f*pi;
        }
        if (GRAD) {
          double *g[3] = {gx, gy, gz};
          #pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            g[ax][2*(size_t)t] += f*rinv*ar[ax];
            g[ax][2*(size_t)t + 1] += f*rinv*ai[ax];
          }
        }
      }
    }
  }
}

// {{{ M2P, polynomial form (compile-time order, potential only)

// Stored multi-indices have m_x = a <= 1, so d^(a,b,c) H = x^a sum_{j,l}
// h(b,j) h(c,l) y^{2j-b} z^{2l-c} g_{a+j+l}, and the whole list-3 sum of one
// source box regroups into
//   sum_s M_s d^s H = sum_a x^a sum_n g_{a+n} R_{a,n}(y, z),
//   R_{a,n} = sum_{beta + gamma <= n} C[a,n,beta,gamma] y^beta z^gamma,
//   C[a,n,beta,gamma] = sum_{j=beta}^{n-gamma} M_(a, 2j-beta, 2(n-j)-gamma)
//                       h(2j-beta, j) h(2(n-j)-gamma, n-j)   (|m| <= p only).
// The coefficients depend on the multipole only: the wave forms them once per
// source box in LDS, then every lane evaluates the polynomials at its target
// by Horner's rule with compile-time offsets (broadcast LDS reads).
// Coefficient order: a, then n, then beta, then gamma.

__host__ __device__ constexpr int tet(int n) { return n*(n + 1)*(n + 2)/6; }

// 2-vector of T (the complex terms of the FP32 list-3/4 forms)
template <typename T>
struct V2T { typedef double2 type; };
template <>
struct V2T<float> { typedef float2 type; };
template <typename T>
__device__ __forceinline__ typename V2T<T>::type mk2(T a, T b) { return {a, b}; }

template <int ORD>
struct M2PPoly {
  static constexpr int N0 = tet(ORD + 1);   // coefficients with a = 0
  static constexpr int NC = N0 + tet(ORD);  // all
};

// stored index of (a, b, c) in the PDE-compressed order (degree d = a+b+c:
// first a = 0 with b = 0..d, then a = 1 with b = 0..d-1)
__device__ __forceinline__ int stored_index(int a, int b, int c) {
  const int d = a + b + c;
  return d*d + (a ? d + 1 + b : b);
}

template <int ORD, typename T>
__device__ __forceinline__ void m2p_poly_eval(const typename V2T<T>::type *C, T x, T y, T z,
                                              const T *gr, const T *gi, T &pr, T &pi) {
  #pragma unroll
  for (int a = 0; a <= 1; ++a) {
    const int base = a ? M2PPoly<ORD>::N0 : 0;
    T ar = 0, ai = 0;
    #pragma unroll
    for (int n = 0; n <= ORD - a; ++n) {
      T Rr = 0, Ri = 0;
      #pragma unroll
      for (int b = n; b >= 0; --b) {
        T Sr = 0, Si = 0;
        #pragma unroll
        for (int c = n - b; c >= 0; --c) {
          const auto v = C[base + tet(n) + b*(n + 1) - b*(b - 1)/2 + c];
          Sr = Sr*z + v.x;
          Si = Si*z + v.y;
        }
        Rr = Rr*y + Sr;
        Ri = Ri*y + Si;
      }
      ar += gr[a + n]*Rr - gi[a + n]*Ri;
      ai += gr[a + n]*Ri + gi[a + n]*Rr;
    }
    if (a == 0) { pr += ar; pi += ai; } else { pr += x*ar; pi += x*ai; }
  }
}

// T = float: the polynomial, its coefficients and g_q in FP32 (g_q from FP64
// values), each source box's term added to FP64 sums
template <int ORD, typename T>
__global__ __launch_bounds__(64) void m2p_helm3d_poly(
    double k, int ntgt, const int *target_boxes, const int *starts, const int *lists,
    const int *bts, const int *btc, const double *centers, int nboxes, const double *tx,
    const double *ty, const double *tz, double rscale, double scale, const double2 *src,
    int sbase, double2 *pot) {
  constexpr int NC = M2PPoly<ORD>::NC, NS = (ORD + 1)*(ORD + 1), HS = ORD + 1;
  __shared__ typename V2T<T>::type sC[NC];
  __shared__ double2 sM[NS];
  __shared__ double ht[HS*HS];
  __shared__ unsigned dec[NC];
  herm_table(ht, ORD);
  const int lane = threadIdx.x;
  for (int i = lane; i < NC; i += 64) {  // (a, n, beta, gamma) of coefficient i
    const int a = i >= M2PPoly<ORD>::N0, r = a ? i - M2PPoly<ORD>::N0 : i;
    int n = 0;
    while (tet(n + 1) <= r) ++n;
    int q = r - tet(n), b = 0;
    while (q >= n + 1 - b) { q -= n + 1 - b; ++b; }
    dec[i] = a | (n << 1) | (b << 8) | (q << 16);
  }
  const double rinv = 1.0/rscale, kappa = k*rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    const int t0 = bts[tb], nt = btc[tb];
    for (int tc = 0; tc < nt; tc += 64) {
      const bool valid = tc + lane < nt;
      const int t = t0 + tc + (valid ? lane : 0);
      const double X = tx[t], Y = ty[t], Z = tz[t];
      double pr = 0, pi = 0;
      for (int qq = starts[ib]; qq < starts[ib + 1]; ++qq) {
        const int sb = lists[qq];
        __syncthreads();  // previous source box done with sC / sM
        const double2 *M = src + (size_t)(sb - sbase)*NS;
        for (int u = lane; u < NS; u += 64) sM[u] = M[u];
        __syncthreads();
        for (int i = lane; i < NC; i += 64) {
          const unsigned d = dec[i];
          const int a = d & 1, n = (d >> 1) & 127, be = (d >> 8) & 255, ga = d >> 16;
          double vr = 0, vi = 0;
          for (int j = be; j <= n - ga; ++j) {
            const int b = 2*j - be, c = 2*(n - j) - ga;
            if (a + b + c > ORD) continue;
            const double w = ht[b*HS + j]*ht[c*HS + (n - j)];
            const double2 m = sM[stored_index(a, b, c)];
            vr += w*m.x;
            vi += w*m.y;
          }
          sC[i] = mk2<T>(T(vr), T(vi));
        }
        const double x = (X - centers[sb])*rinv, y = (Y - centers[nboxes + sb])*rinv,
                     z = (Z - centers[2*(size_t)nboxes + sb])*rinv;
        double gr[ORD + 1], gi[ORD + 1];
        {
          const double r = sqrt(x*x + y*y + z*z), ri = 1.0/r, ri2 = ri*ri;
          double sn, cs;
          sincos(kappa*r, &sn, &cs);
          gr[0] = cs*ri;
          gi[0] = sn*ri;
          if (ORD >= 1) {
            const double kr = kappa*r;
            gr[1] = (-cs - sn*kr)*ri2*ri;
            gi[1] = (cs*kr - sn)*ri2*ri;
          }
          #pragma unroll
          for (int q = 2; q <= ORD; ++q) {
            gr[q] = (-(2*q - 1)*gr[q - 1] - kappa*kappa*gr[q - 2])*ri2;
            gi[q] = (-(2*q - 1)*gi[q - 1] - kappa*kappa*gi[q - 2])*ri2;
          }
        }
        __syncthreads();
        if constexpr (sizeof(T) == sizeof(double)) {
          m2p_poly_eval<ORD, double>(sC, x, y, z, gr, gi, pr, pi);
        } else {
          T grt[ORD + 1], git[ORD + 1];
          #pragma unroll
          for (int q = 0; q <= ORD; ++q) { grt[q] = T(gr[q]); git[q] = T(gi[q]); }
          T br = 0, bi = 0;
          m2p_poly_eval<ORD, T>(sC, T(x), T(y), T(z), grt, git, br, bi);
          pr += double(br);
          pi += double(bi);
        }
      }
      if (valid) {
        pot[t].x += scale*rinv*pr;
        pot[t].y += scale*rinv*pi;
      }
    }
  }
}

// }}}

// {{{ P2L

constexpr int P2LH_CH = 64;  // sources staged per chunk

// One block per target box; thread = stored local coefficient t.  The list-4
// sources are staged in chunks: per source the powers of xi = (c - y)/rho,
// g_q and the weights; then
//   L_t = sum_y [w0 rho^|t| d^t G(c - y) + sum_i w_i rho^|t| d^{t + e_i} G(c - y)/rho]
// (reference sumpy/p2e.py:268-388; overwrites the box's row).
template <bool DIP>
__global__ __launch_bounds__(256) void p2l_helm3d(
    sumpyhip_expn e, double k, double kim, int ntgt, const int *target_boxes, const int *starts,
    const int *lists, const int *bss, const int *bsc, const double *centers, int nboxes,
    const double *sx, const double *sy, const double *sz, const double *w0,
    const double *wx, const double *wy, const double *wz, double rscale, double *tgt,
    int tbase) {
  extern __shared__ double sh[];
  const int p = e.order, nst = e.nstored;
  const int NQ = p + (DIP ? 1 : 0);
  const int tid = threadIdx.x;
  double *ht = sh;                                  // [(NQ+1)^2]
  double *tab = ht + (NQ + 1)*(NQ + 1);             // [5*(NQ+1)][CH]
  double *sw = tab + 5*(NQ + 1)*P2LH_CH;            // [(DIP ? 8 : 2)][CH]
  herm_table(ht, NQ);
  const double *xp = tab + 0*(NQ + 1)*P2LH_CH, *yp = tab + 1*(NQ + 1)*P2LH_CH,
               *zp = tab + 2*(NQ + 1)*P2LH_CH, *gre = tab + 3*(NQ + 1)*P2LH_CH,
               *gim = tab + 4*(NQ + 1)*P2LH_CH;
  const bool own = tid < nst;
  const int a = own ? e.stored_ids[tid] : 0, b = own ? e.stored_ids[nst + tid] : 0,
            c = own ? e.stored_ids[2*nst + tid] : 0;
  const double rinv = 1.0/rscale, kappa = k*rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    const double cx = centers[tb], cy = centers[nboxes + tb],
                 cz = centers[2*(size_t)nboxes + tb];
    double lr = 0, li = 0;
    for (int q = starts[ib]; q < starts[ib + 1]; ++q) {
      const int sbox = lists[q];
      const int s0 = bss[sbox], ns = bsc[sbox];
      for (int sc = 0; sc < ns; sc += P2LH_CH) {
        const int n = min(P2LH_CH, ns - sc);
        __syncthreads();
        if (tid < n) {
          const int s = s0 + sc + tid;
          const double x = (cx - sx[s])*rinv, y = (cy - sy[s])*rinv, z = (cz - sz[s])*rinv;
          double px = 1, py = 1, pz = 1;
          for (int u = 0; u <= NQ; ++u) {
            tab[(0*(NQ + 1) + u)*P2LH_CH + tid] = px;
            tab[(1*(NQ + 1) + u)*P2LH_CH + tid] = py;
            tab[(2*(NQ + 1) + u)*P2LH_CH + tid] = pz;
            px *= x; py *= y; pz *= z;
          }
          helm_radial(kappa, kim*rscale, sqrt(x*x + y*y + z*z), NQ,
                      tab + 3*(NQ + 1)*P2LH_CH + tid,
                      tab + 4*(NQ + 1)*P2LH_CH + tid, P2LH_CH);
          sw[tid] = w0[2*(size_t)s];
          sw[P2LH_CH + tid] = w0[2*(size_t)s + 1];
          if (DIP) {
            const double *wd[3] = {wx, wy, wz};
            #pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
              sw[(2 + 2*ax)*P2LH_CH + tid] = wd[ax][2*(size_t)s];
              sw[(3 + 2*ax)*P2LH_CH + tid] = wd[ax][2*(size_t)s + 1];
            }
          }
        }
        __syncthreads();
        if (own) {
          for (int j = 0; j < n; ++j) {
            const cplx d = helm_deriv(a, b, c, xp + j, yp + j, zp + j, P2LH_CH, gre + j,
                                      gim + j, P2LH_CH, ht, NQ + 1);
            const double qr = sw[j], qi = sw[P2LH_CH + j];
            lr += qr*d.re - qi*d.im;
            li += qr*d.im + qi*d.re;
            if (DIP) {
              #pragma unroll
              for (int ax = 0; ax < 3; ++ax) {
                const cplx g = helm_deriv(a + (ax == 0), b + (ax == 1), c + (ax == 2),
                                          xp + j, yp + j, zp + j, P2LH_CH, gre + j, gim + j,
                                          P2LH_CH, ht, NQ + 1);
                const double ur = sw[(2 + 2*ax)*P2LH_CH + j]*rinv,
                             ui = sw[(3 + 2*ax)*P2LH_CH + j]*rinv;
                lr += ur*g.re - ui*g.im;
                li += ur*g.im + ui*g.re;
              }
            }
          }
        }
      }
    }
    if (own) {
      double *o = tgt + 2*((size_t)(tb - tbase)*nst + tid);
      o[0] = lr*rinv;
      o[1] = li*rinv;
    }
  }
}

// }}}

// {{{ P2L, moment form (compile-time order, charges only)

// The transpose of the polynomial M2P: with xi = (c - y)/rho,
//   L_(a,b,c) = sum_{j,l} h(b,j) h(c,l) Phi_a(j + l, 2j - b, 2l - c),
//   Phi_a(n, beta, gamma) = sum_y q_y g_{a+n}(|xi_y|) x_y^a y_y^beta z_y^gamma,
// (beta + gamma <= n <= p - a), then times 1/rho.  Thread r owns the moment row
// (a, n, beta) -- (p+1)^2 rows, gamma = 0..n-beta -- in registers while the
// box's list-4 sources stream through LDS; the coefficients are assembled
// from the moments at the end.

// T = float: the staged terms (q g_q, powers) and each chunk's moment sums
// in FP32, added to FP64 totals per chunk of CH sources
template <int ORD, typename T>
__global__ __launch_bounds__(256) void p2l_helm3d_moments(
    double k, int ntgt, const int *target_boxes, const int *starts, const int *lists,
    const int *bss, const int *bsc, const double *centers, int nboxes, const double *sx,
    const double *sy, const double *sz, const double2 *w0, double rscale, double2 *tgt,
    int tbase) {
  constexpr int NS = (ORD + 1)*(ORD + 1), NC = M2PPoly<ORD>::NC, HS = ORD + 1;
  constexpr int CH = P2LH_CH, NP = ORD + 1;
  // the source staging (q g_q, powers) and, after the last source, the
  // moments share one LDS buffer
  typedef typename V2T<T>::type V2;
  constexpr int STAGE = NP*CH*2 + 2*NP*CH + CH;  // doubles
  constexpr int BUF = STAGE > 2*NC ? STAGE : 2*NC;
  __shared__ __align__(16) double sbuf[BUF];
  V2 (*sQG)[CH] = reinterpret_cast<V2 (*)[CH]>(sbuf);
  T (*sY)[CH] = reinterpret_cast<T (*)[CH]>(sbuf + NP*CH*2);
  T (*sZ)[CH] = reinterpret_cast<T (*)[CH]>(sbuf + NP*CH*3);
  T *sX = reinterpret_cast<T *>(sbuf + NP*CH*4);
  double2 *sPhi = reinterpret_cast<double2 *>(sbuf);
  __shared__ double ht[HS*HS];
  herm_table(ht, ORD);
  const int tid = threadIdx.x;
  // moment row of this thread
  int ra = 0, rn = 0, rb = 0, rlen = 0, rbase = 0;
  if (tid < NS) {
    constexpr int R0 = (ORD + 1)*(ORD + 2)/2;
    ra = tid >= R0;
    int r = ra ? tid - R0 : tid;
    while (r > rn) { r -= rn + 1; ++rn; }
    rb = r;
    rlen = rn - rb + 1;
    rbase = (ra ? M2PPoly<ORD>::N0 : 0) + tet(rn) + rb*(rn + 1) - rb*(rb - 1)/2;
  }
  const double rinv = 1.0/rscale, kappa = k*rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    const double cx = centers[tb], cy = centers[nboxes + tb],
                 cz = centers[2*(size_t)nboxes + tb];
    double pr[NP], pi[NP];
    #pragma unroll
    for (int g = 0; g < NP; ++g) { pr[g] = 0; pi[g] = 0; }
    for (int q = starts[ib]; q < starts[ib + 1]; ++q) {
      const int sbox = lists[q];
      const int s0 = bss[sbox], ns = bsc[sbox];
      for (int sc = 0; sc < ns; sc += CH) {
        const int n = min(CH, ns - sc);
        __syncthreads();
        if (tid < n) {
          const int si = s0 + sc + tid;
          const double x = (cx - sx[si])*rinv, y = (cy - sy[si])*rinv, z = (cz - sz[si])*rinv;
          const double r = sqrt(x*x + y*y + z*z), ri = 1.0/r, ri2 = ri*ri;
          double sn, cs;
          sincos(kappa*r, &sn, &cs);
          const double2 w = w0[si];
          double ar = cs*ri, ai = sn*ri;
          sQG[0][tid] = mk2<T>(T(w.x*ar - w.y*ai), T(w.x*ai + w.y*ar));
          double br = 0, bi = 0;
          if (ORD >= 1) {
            const double kr = kappa*r;
            br = (-cs - sn*kr)*ri2*ri;
            bi = (cs*kr - sn)*ri2*ri;
            sQG[1][tid] = mk2<T>(T(w.x*br - w.y*bi), T(w.x*bi + w.y*br));
          }
          #pragma unroll
          for (int g = 2; g <= ORD; ++g) {
            const double f = -(2*g - 1);
            const double cr = (f*br - kappa*kappa*ar)*ri2, ci = (f*bi - kappa*kappa*ai)*ri2;
            sQG[g][tid] = mk2<T>(T(w.x*cr - w.y*ci), T(w.x*ci + w.y*cr));
            ar = br; ai = bi; br = cr; bi = ci;
          }
          double py = 1, pz = 1;
          #pragma unroll
          for (int e = 0; e <= ORD; ++e) {
            sY[e][tid] = T(py);
            sZ[e][tid] = T(pz);
            py *= y;
            pz *= z;
          }
          sX[tid] = T(x);
        }
        __syncthreads();
        if constexpr (sizeof(T) == sizeof(double)) {
          if (tid < NS) {
            #pragma unroll 2
            for (int j = 0; j < n; ++j) {
              const V2 qg = sQG[ra + rn][j];
              const double m = (ra ? sX[j] : 1.0)*sY[rb][j];
              const double wr = qg.x*m, wi = qg.y*m;
              #pragma unroll
              for (int g = 0; g < NP; ++g) {
                if (g < rlen) {
                  const double zg = sZ[g][j];
                  pr[g] += wr*zg;
                  pi[g] += wi*zg;
                }
              }
            }
          }
        } else if (tid < NS) {
          T cr[NP], ci[NP];
          #pragma unroll
          for (int g = 0; g < NP; ++g) { cr[g] = 0; ci[g] = 0; }
          #pragma unroll 2
          for (int j = 0; j < n; ++j) {
            const V2 qg = sQG[ra + rn][j];
            const T m = (ra ? sX[j] : T(1))*sY[rb][j];
            const T wr = qg.x*m, wi = qg.y*m;
            #pragma unroll
            for (int g = 0; g < NP; ++g) {
              if (g < rlen) {
                const T zg = sZ[g][j];
                cr[g] += wr*zg;
                ci[g] += wi*zg;
              }
            }
          }
          #pragma unroll
          for (int g = 0; g < NP; ++g) { pr[g] += double(cr[g]); pi[g] += double(ci[g]); }
        }
      }
    }
    __syncthreads();
    if (tid < NS) {
      #pragma unroll
      for (int g = 0; g < NP; ++g)
        if (g < rlen) sPhi[rbase + g] = make_double2(pr[g], pi[g]);
    }
    __syncthreads();
    if (tid < NS) {
      // coefficient tid = (a, b, c) in the stored order
      int d = 0;
      while ((d + 1)*(d + 1) <= tid) ++d;
      const int o = tid - d*d, a = o > d, b = a ? o - d - 1 : o, c = d - a - b;
      const int base = a ? M2PPoly<ORD>::N0 : 0;
      double lr = 0, li = 0;
      for (int j = (b + 1)/2; j <= b; ++j)
        for (int l = (c + 1)/2; l <= c; ++l) {
          const int nn = j + l, be = 2*j - b, ga = 2*l - c;
          const double2 v = sPhi[base + tet(nn) + be*(nn + 1) - be*(be - 1)/2 + ga];
          const double f = ht[b*HS + j]*ht[c*HS + l];
          lr += f*v.x;
          li += f*v.y;
        }
      tgt[(size_t)(tb - tbase)*NS + tid] = make_double2(lr*rinv, li*rinv);
    }
  }
}

// }}}

// {{{ M2L without translation classes

// One block per target box; per list-2 source box the block writes the
// scaled derivatives at the pair's vector (as the class-data kernel does for
// a class) into LDS, then thread t adds sum_s M_s D(t + s) (complex) to its
// local coefficient; the sums are scaled by (tgt_rscale/src_rscale)^|t| and
// overwrite the target rows (reference sumpy/e2e.py:164-304).
__global__ __launch_bounds__(256) void m2l_csr_helm3d(
    int N, int M0, int M1, int M2, int nl, int nm, const int *off_t, const int *off_s,
    const int *deg, int ntgt, const int *target_boxes, const int *starts, const int *lists,
    const double *centers, int nboxes, double src_rscale, double tgt_rscale, double k,
    double kim, const double2 *src, int sbase, double2 *tgt, int tbase) {
  extern __shared__ double lds[];
  const int ncirc = M0*M1*M2, HS = N + 1;
  double *Dre = lds, *Dim = Dre + ncirc;
  double2 *Ms = reinterpret_cast<double2 *>(Dim + ncirc);  // 2 ncirc doubles: 16-byte aligned
  double *ht = reinterpret_cast<double *>(Ms + nm);
  double *sp = ht + HS*HS, *sg = sp + 3*HS;
  herm_table(ht, N);
  const int tid = threadIdx.x;
  const double rinv = 1.0/src_rscale, ratio = tgt_rscale/src_rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
    for (int is = starts[ib]; is < starts[ib + 1]; ++is) {
      const int sb = lists[is];
      __syncthreads();  // the previous pair is done with D, Ms, sp, sg
      for (int q = tid; q < nm; q += 256) Ms[q] = src[(size_t)(sb - sbase)*nm + q];
      if (tid < 4) {
        const double x = (centers[tb] - centers[sb])*rinv,
                     y = (centers[nboxes + tb] - centers[nboxes + sb])*rinv,
                     z = (centers[2*(size_t)nboxes + tb] - centers[2*(size_t)nboxes + sb])*rinv;
        if (tid < 3) {
          const double xi = tid == 0 ? x : (tid == 1 ? y : z);
          double pw = 1;
          for (int e = 0; e <= N; ++e) {
            sp[tid*HS + e] = pw;
            pw *= xi;
          }
        } else {
          helm_radial(k*src_rscale, kim*src_rscale, sqrt(x*x + y*y + z*z), N, sg, sg + HS, 1);
        }
      }
      __syncthreads();
      for (int e = tid; e < ncirc; e += 256) {
        const int mz = e % M2, my = (e/M2) % M1, mx = e/(M2*M1);
        double re = 0, im = 0;
        if (mx + my + mz <= N) {
          const cplx d = helm_deriv(mx, my, mz, sp, sp + HS, sp + 2*HS, 1, sg, sg + HS, 1,
                                    ht, HS);
          re = d.re*rinv;
          im = d.im*rinv;
        }
        Dre[e] = re;
        Dim[e] = im;
      }
      __syncthreads();
      for (int q = 0; q < 4; ++q) {
        const int row = tid + 256*q;
        if (row >= nl) break;
        const int ot = off_t[row];
        double sr = 0, si = 0;
        for (int kk = 0; kk < nm; ++kk) {
          const double2 m = Ms[kk];
          const double dr = Dre[ot + off_s[kk]], di = Dim[ot + off_s[kk]];
          sr += m.x*dr - m.y*di;
          si += m.x*di + m.y*dr;
        }
        ar[q] += sr;
        ai[q] += si;
      }
    }
    for (int q = 0; q < 4; ++q) {
      const int row = tid + 256*q;
      if (row >= nl) break;
      const double f = pow(ratio, (double)deg[row]);
      tgt[(size_t)(tb - tbase)*nl + row] = make_double2(ar[q]*f, ai[q]*f);
    }
  }
}

// }}}

int check_helm_expn(const sumpyhip_expn *e, const char *what) {
  SUMPY_REQUIRE(e != nullptr && e->stored_ids != nullptr, "%s: null expansion tables", what);
  SUMPY_REQUIRE(e->order >= 0 && e->order + 2 <= HELM_MAXN, "%s: order %d out of range",
                what, e->order);
  SUMPY_REQUIRE(e->nstored > 0 && e->nstored <= 256, "%s: %d stored coefficients (at most 256)",
                what, e->nstored);
  return SUMPYHIP_OK;
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" int sumpyhip_m2l_class_data_helmholtz(int order_src, int order_tgt,
                                                 const int *circ_shape, int nclasses,
                                                 const double *translation_vectors,
                                                 int nvectors, int class_start,
                                                 double src_rscale, double k, double k_im,
                                                 double *out, void *stream) {
  SUMPY_REQUIRE(circ_shape != nullptr, "m2l_class_data_helmholtz: null shape");
  const int N = order_src + order_tgt;
  SUMPY_REQUIRE(N >= 0 && N <= HELM_MAXN, "m2l_class_data_helmholtz: order %d > %d", N,
                HELM_MAXN);
  SUMPY_REQUIRE(src_rscale > 0, "m2l_class_data_helmholtz: rscale must be positive");
  if (nclasses == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(class_start >= 0 && class_start + nclasses <= nvectors,
                "m2l_class_data_helmholtz: class range outside the vectors");
  const int grid = nclasses < 1024 ? nclasses : 1024;
  helm_class_data_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(
      N, circ_shape[0], circ_shape[1], circ_shape[2], nclasses, translation_vectors,
      nvectors, class_start, src_rscale, k, k_im, out);
  return check_launch("m2l_class_data_helmholtz");
}

extern "C" int sumpyhip_m2p_helmholtz(const sumpyhip_expn *expn, double k, double k_im,
                                      int ntgt_boxes,
                                      const int *target_boxes, const int *source_box_starts,
                                      const int *source_box_lists,
                                      const int *box_target_starts,
                                      const int *box_target_counts_nonchild,
                                      const double *box_centers, int nboxes,
                                      const double *tx, const double *ty, const double *tz,
                                      double rscale, double scale,
                                      const void *src_expansions, int src_base_ibox,
                                      void *pot, void *gx, void *gy, void *gz,
                                      int precision, void *stream) {
  int rc = check_helm_expn(expn, "m2p_helmholtz");
  SUMPY_REQUIRE(precision == 0 || precision == 1, "m2p_helmholtz: precision %d", precision);
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(pot || gx, "m2p_helmholtz: no output");
  SUMPY_REQUIRE(!gx || (gy && gz), "m2p_helmholtz: incomplete gradient outputs");
  const int NQ = expn->order + 1 + (gx ? 1 : 0);
  const size_t lds = sizeof(double)*((NQ + 1)*(NQ + 1) + 2*expn->nstored + 5*(NQ + 1)*64) +
                     sizeof(int)*3*expn->nstored;
  SUMPY_REQUIRE(lds <= 64*1024, "m2p_helmholtz: order too large");
  const int grid = ntgt_boxes < 65536 ? ntgt_boxes : 65536;
  hipStream_t st = (hipStream_t)stream;
  const double *src = static_cast<const double *>(src_expansions);
  if (k_im == 0.0 && !gx && pot && expn->nstored == (expn->order + 1)*(expn->order + 1)) {
    // potentials of the usual orders: the polynomial form
#define SUMPY_M2P_POLY(O)                                                                  \
  case O:                                                                                  \
    if (precision)                                                                         \
      m2p_helm3d_poly<O, float><<<grid, 64, 0, st>>>(                                      \
          k, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,                \
          box_target_starts, box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz,  \
          rscale, scale, static_cast<const double2 *>(src_expansions), src_base_ibox,      \
          static_cast<double2 *>(pot));                                                    \
    else                                                                                   \
      m2p_helm3d_poly<O, double><<<grid, 64, 0, st>>>(                                     \
          k, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,                \
          box_target_starts, box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz,  \
          rscale, scale, static_cast<const double2 *>(src_expansions), src_base_ibox,      \
          static_cast<double2 *>(pot));                                                    \
    return check_launch("m2p_helmholtz");
    switch (expn->order) {
      SUMPY_M2P_POLY(1) SUMPY_M2P_POLY(2) SUMPY_M2P_POLY(3) SUMPY_M2P_POLY(4)
      SUMPY_M2P_POLY(5) SUMPY_M2P_POLY(6) SUMPY_M2P_POLY(7) SUMPY_M2P_POLY(8)
      SUMPY_M2P_POLY(9) SUMPY_M2P_POLY(10) SUMPY_M2P_POLY(11) SUMPY_M2P_POLY(12)
      default: break;
    }
#undef SUMPY_M2P_POLY
  }
  if (gx)
    m2p_helm3d<true><<<grid, 64, lds, st>>>(
        *expn, k, k_im, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,
        box_target_starts, box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz,
        rscale, scale, src, src_base_ibox, (double *)pot, (double *)gx, (double *)gy,
        (double *)gz);
  else
    m2p_helm3d<false><<<grid, 64, lds, st>>>(
        *expn, k, k_im, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,
        box_target_starts, box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz,
        rscale, scale, src, src_base_ibox, (double *)pot, nullptr, nullptr, nullptr);
  return check_launch("m2p_helmholtz");
}

extern "C" int sumpyhip_p2l_helmholtz(const sumpyhip_expn *expn, double k, double k_im,
                                      int ntgt_boxes,
                                      const int *target_boxes, const int *source_box_starts,
                                      const int *source_box_lists,
                                      const int *box_source_starts,
                                      const int *box_source_counts_nonchild,
                                      const double *box_centers, int nboxes,
                                      const double *sx, const double *sy, const double *sz,
                                      const void *w0, const void *wx, const void *wy,
                                      const void *wz, double rscale, void *tgt_expansions,
                                      int tgt_base_ibox, int precision, void *stream) {
  int rc = check_helm_expn(expn, "p2l_helmholtz");
  SUMPY_REQUIRE(precision == 0 || precision == 1, "p2l_helmholtz: precision %d", precision);
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(w0 != nullptr, "p2l_helmholtz: null charge weights");
  SUMPY_REQUIRE(!wx || (wy && wz), "p2l_helmholtz: incomplete dipole weights");
  const int NQ = expn->order + (wx ? 1 : 0);
  const size_t lds = sizeof(double)*((NQ + 1)*(NQ + 1) + 5*(NQ + 1)*P2LH_CH + 8*P2LH_CH);
  SUMPY_REQUIRE(lds <= 64*1024, "p2l_helmholtz: order too large");
  const int nt = (expn->nstored + 63)/64*64;
  const int grid = ntgt_boxes < 65536 ? ntgt_boxes : 65536;
  hipStream_t st = (hipStream_t)stream;
  const double *c0 = static_cast<const double *>(w0);
  if (k_im == 0.0 && !wx && expn->nstored == (expn->order + 1)*(expn->order + 1)) {
    // charges at the usual orders: the moment form
#define SUMPY_P2L_MOM(O)                                                                   \
  case O:                                                                                  \
    if (precision)                                                                         \
      p2l_helm3d_moments<O, float><<<grid, nt, 0, st>>>(                                   \
          k, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,                \
          box_source_starts, box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz,  \
          static_cast<const double2 *>(w0), rscale,                                        \
          static_cast<double2 *>(tgt_expansions), tgt_base_ibox);                          \
    else                                                                                   \
      p2l_helm3d_moments<O, double><<<grid, nt, 0, st>>>(                                  \
          k, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,                \
          box_source_starts, box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz,  \
          static_cast<const double2 *>(w0), rscale,                                        \
          static_cast<double2 *>(tgt_expansions), tgt_base_ibox);                          \
    return check_launch("p2l_helmholtz");
    switch (expn->order) {
      SUMPY_P2L_MOM(1) SUMPY_P2L_MOM(2) SUMPY_P2L_MOM(3) SUMPY_P2L_MOM(4)
      SUMPY_P2L_MOM(5) SUMPY_P2L_MOM(6) SUMPY_P2L_MOM(7) SUMPY_P2L_MOM(8)
      SUMPY_P2L_MOM(9) SUMPY_P2L_MOM(10) SUMPY_P2L_MOM(11) SUMPY_P2L_MOM(12)
      default: break;
    }
#undef SUMPY_P2L_MOM
  }
  if (wx)
    p2l_helm3d<true><<<grid, nt, lds, st>>>(
        *expn, k, k_im, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,
        box_source_starts, box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz, c0,
        (const double *)wx, (const double *)wy, (const double *)wz, rscale,
        (double *)tgt_expansions, tgt_base_ibox);
  else
    p2l_helm3d<false><<<grid, nt, lds, st>>>(
        *expn, k, k_im, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,
        box_source_starts, box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz, c0,
        nullptr, nullptr, nullptr, rscale, (double *)tgt_expansions, tgt_base_ibox);
  return check_launch("p2l_helmholtz");
}

extern "C" int sumpyhip_m2l_csr_helmholtz(int order_src, int order_tgt, const int *circ_shape,
                                          int nl, int nm, const int *off_t, const int *off_s,
                                          const int *tgt_degree, int ntgt_boxes,
                                          const int *target_boxes, const int *src_box_starts,
                                          const int *src_box_lists, const double *box_centers,
                                          int nboxes, double src_rscale, double tgt_rscale,
                                          double k, double k_im, const void *src_expansions,
                                          int src_base_ibox, void *tgt_expansions,
                                          int tgt_base_ibox, void *stream) {
  SUMPY_REQUIRE(circ_shape != nullptr, "m2l_csr_helmholtz: null shape");
  const int N = order_src + order_tgt;
  SUMPY_REQUIRE(N >= 0 && N <= HELM_MAXN, "m2l_csr_helmholtz: order %d > %d", N, HELM_MAXN);
  SUMPY_REQUIRE(nl <= 1024, "m2l_csr_helmholtz: too many target coefficients");
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  const size_t ncirc = (size_t)circ_shape[0]*circ_shape[1]*circ_shape[2];
  const size_t HS = N + 1;
  const size_t lds = sizeof(double)*(2*ncirc + 2*(size_t)nm + HS*HS + 5*HS);
  SUMPY_REQUIRE(lds <= 64*1024, "m2l_csr_helmholtz: %zu bytes of LDS needed (order too high)",
                lds);
  const int grid = ntgt_boxes < 8192 ? ntgt_boxes : 8192;
  m2l_csr_helm3d<<<grid, 256, lds, (hipStream_t)stream>>>(
      N, circ_shape[0], circ_shape[1], circ_shape[2], nl, nm, off_t, off_s, tgt_degree,
      ntgt_boxes, target_boxes, src_box_starts, src_box_lists, box_centers, nboxes,
      src_rscale, tgt_rscale, k, k_im, static_cast<const double2 *>(src_expansions),
      src_base_ibox,
      static_cast<double2 *>(tgt_expansions), tgt_base_ibox);
  return check_launch("m2l_csr_helmholtz");
}
