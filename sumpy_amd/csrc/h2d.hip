// 2D Helmholtz and Yukawa FMMs with Hankel/Bessel expansions (H2D, Y2D).
//
// Helmholtz: G(d) = (i/4) H0^(1)(k |d|), real k > 0 or complex k with
// Im k > 0 (damped waves; sumpy/kernel.py:754-844, 2D); Yukawa: G(d) = (i/4) H0^(1)(i lam |d|) = K0(lam |d|)/(2 pi), lam > 0
// (sumpy/kernel.py:848-897).  The reference's H2D / Y2D expansions
// (sumpy/expansion/multipole.py:426-567, sumpy/expansion/local.py:522-704)
// and its Fourier-Bessel M2L (sumpy/expansion/m2l.py:1008-1115) store 2p+1
// complex coefficients c = -p..p at index c + p; with the Bessel argument
// scale s = k (Helmholtz) or i lam (Yukawa, multipole.py:552-567):
//   P2M  M_c = sum_s q_s J_c(k|a|) rscale^-|c| e^{-i c th_s}
//   M2P  u(t) = (i/4) sum_c M_c H_c(k|b|) rscale^|c| e^{i c th_b}
//   P2L  L_c = sum_s q_s H_c(k|a|) rscale^|c| e^{i c th_s}
//   L2P  u(t) = (i/4) sum_c L_c J_c(k|b|) rscale^-|c| e^{-i c th_b}
//   M2M  M'_j = sum_m M_m J_{m-j}(k|d|) rs^|m| rt^-|j| e^{i (m-j) th_d}
//   L2L  L'_j = sum_m L_m J_{m-j}(k|d|) rs^-|m| rt^|j| e^{-i (m-j) th_d}
//   M2L  L_j = (-1)^j rt^|j| sum_m D[m + j + ps + pt] rs^|m| M_m,
//        D[n + ps + pt] = H_n(k|d|) e^{i n th_d}  (the translation-class data)
// with a = centre - source, b = target - centre, d = new centre - old
// centre, th_s the angle of the source about the centre, and J_c, H_c taken
// at s |.| (the formulas above are written for s = k).
//
// Bessel functions (KIND 0, Helmholtz): J_0..J_n by Miller's backward
// recurrence normalised with J_0 + 2 sum_k J_2k = 1; Y_0 and Y_1 from their
// Neumann series in the J_n (Abramowitz & Stegun 9.1.88), Y_n by the upward
// recurrence (stable for Y); H_n = J_n + i Y_n.  (KIND 1, Yukawa) with
// x = lam r: J_n(i x) = i^n I_n(x), H_n(i x) = (2/pi) i^(-n-1) K_n(x); I_n by
// Miller's recurrence normalised with I_0 + 2 sum_k I_k = e^x; K_0, K_1 by the
// trapezoidal rule on K_nu(x) = int_0^inf e^{-x cosh t} cosh(nu t) dt (step
// min(1/4, 1/(2 sqrt x)), exponentially convergent for this integrand), K_n
// by the upward recurrence.  (KIND 2, complex k) the same Miller / Neumann
// route in complex arithmetic (bessel_jy_c); H_n = J_n + i Y_n then loses
// about e^{2 Im(k) r} relative to |H_n|, which limits it to moderate Im(k) r.
// All kinds: Z_-n = (-1)^n Z_n for Z = J, H.
//
// Layout: one wave (64 lanes) per box; lane c + p owns coefficient c
// (2p+1 <= 63), so the orders go up to 31.  Per-source terms are summed per
// lane and reduced over the wave with a fixed butterfly (deterministic).
#include "common.h"
#include "bessel.h"

#include <cmath>
#include <cstdint>

namespace sumpyhip {
namespace {

constexpr int PMAX = 31;
constexpr int NSEQ = 2*PMAX + 2;  // sequences of orders 0..2 PMAX (+1 for Y_1)
constexpr int WAVE = 64;

__device__ __forceinline__ cplx conj(cplx a) { return {a.re, -a.im}; }
__device__ __forceinline__ cplx cfma(cplx a, cplx b, cplx acc) {
  return {acc.re + a.re*b.re - a.im*b.im, acc.im + a.re*b.im + a.im*b.re};
}

// The two function sequences of one kernel kind at distance r, as complex
// values Z^J_c = J_c(s r) and Z^H_c = H_c(s r), c = 0..n (n >= 1 for H), with
// the Bessel argument scale s = k (KIND 0, real k; KIND 2, complex k) or
// i lam (KIND 1; kk = (lam, 0)).
template <int KIND>
__device__ __forceinline__ void seq_j(cplx kk, double r, int n, cplx *Z) {
  if (KIND == 2) {
    bessel_jy_c(cmul(kk, cplx{r, 0.0}), n, Z, nullptr);
    return;
  }
  double A[NSEQ];
  if (KIND == 0) bessel_jy(kk.re*r, n, A, nullptr);
  else bessel_ik(kk.re*r, n, A, nullptr);
  for (int c = 0; c <= n; ++c) {
    if (KIND == 0) {
      Z[c] = {A[c], 0.0};
    } else {
      // i^c I_c
      const int q = c & 3;
      Z[c] = q == 0 ? cplx{A[c], 0.0} : q == 1 ? cplx{0.0, A[c]}
                    : q == 2 ? cplx{-A[c], 0.0} : cplx{0.0, -A[c]};
    }
  }
}

template <int KIND>
__device__ __forceinline__ void seq_h(cplx kk, double r, int n, cplx *Z) {
  if (KIND == 2) {
    cplx Y[NSEQ];
    bessel_jy_c(cmul(kk, cplx{r, 0.0}), n, Z, Y);
    for (int c = 0; c <= n; ++c) Z[c] = {Z[c].re - Y[c].im, Z[c].im + Y[c].re};  // J + i Y
    return;
  }
  double A[NSEQ], B[NSEQ];
  if (KIND == 0) bessel_jy(kk.re*r, n, A, B);
  else bessel_ik(kk.re*r, n, A, B);
  for (int c = 0; c <= n; ++c) {
    if (KIND == 0) {
      Z[c] = {A[c], B[c]};
    } else {
      // (2/pi) i^(-c-1) K_c
      const double v = (2.0/PI)*B[c];
      const int q = c & 3;
      Z[c] = q == 0 ? cplx{0.0, -v} : q == 1 ? cplx{-v, 0.0}
                    : q == 2 ? cplx{0.0, v} : cplx{v, 0.0};
    }
  }
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum(v, WAVE); }

__device__ __forceinline__ cplx load_c(const double *p, long i) { return {p[2*i], p[2*i + 1]}; }
__device__ __forceinline__ void store_c(double *p, long i, cplx v) {
  p[2*i] = v.re;
  p[2*i + 1] = v.im;
}

// rscale^e for 0 <= e <= 2 PMAX (e multiplications: the same on every lane
// for one e, no pow)
__device__ __forceinline__ double ipow(double r, int e) {
  double v = 1.0;
  for (int i = 0; i < e; ++i) v *= r;
  return v;
}

// unit vector of v (1, 0 for v = 0) and |v|
__device__ __forceinline__ double unit(double vx, double vy, double &ux, double &uy) {
  const double r = sqrt(vx*vx + vy*vy);
  if (r > 0.0) {
    ux = vx/r;
    uy = vy/r;
  } else {
    ux = 1.0;
    uy = 0.0;
  }
  return r;
}

// Per-lane accumulation of sum_c w Z_c e^{i s c th} over c = -p..p into
// acc[c + p] (Z: the sequence, c >= 0); the phase e^{i th} is (ux, uy) and
// s = sign.
__device__ void add_terms(int p, cplx w, const cplx *Z, double ux, double uy, int sign,
                          cplx *acc) {
  const cplx e1 = {ux, sign*uy};
  cplx ph = {1.0, 0.0};  // e^{i s c th}
  for (int c = 0; c <= p; ++c) {
    const cplx z = Z[c];
    const cplx t = cmul(w, cmul(z, ph));
    acc[p + c] = cadd(acc[p + c], t);
    if (c > 0) {
      // c -> -c: Z_-c = (-1)^c Z_c, e^{-i s c th} = conj(ph)
      const cplx tn = cmul(w, cmul(z, conj(ph)));
      acc[p - c] = cadd(acc[p - c], (c & 1) ? cscale(tn, -1.0) : tn);
    }
    ph = cmul(ph, e1);
  }
}

// Bessel argument scale s: k (Helmholtz, real or complex) or i lam (Yukawa)
template <int KIND>
__device__ __forceinline__ cplx sarg(cplx kk) {
  return KIND == 1 ? cplx{0.0, kk.re} : kk;
}

// acc[p + c] += v for |c| <= p
__device__ __forceinline__ void acc_add(int p, int c, cplx v, cplx *acc) {
  if (c >= -p && c <= p) acc[p + c] = cadd(acc[p + c], v);
}

// Derivatives of F_c = Z_c(s r) e^{i sg c th} (sg = +-1) by the shifts
// W F_c = -s F_{c+1}, Wbar F_c = s F_{c-1} (W = d_x + i d_y):
//   d_x F_c = (s/2) (F_{c-1} - F_{c+1}),  d_y F_c = (i sg s/2) (F_{c-1} + F_{c+1}),
// so n . grad F_c = (s/2) ((n_x + i sg n_y) F_{c-1} - (n_x - i sg n_y) F_{c+1}).
// Per-lane accumulation of w F_c + a F_{c-1} - b F_{c+1} over c = -p..p; Z
// holds Z_0..Z_{p+1}.  A dipole weight u (contributing u . grad_d G(t - s),
// i.e. -u . grad_s) gives a = -(s/2)(u_x + i sg u_y), b = -(s/2)(u_x - i sg u_y).
__device__ void add_terms_der(int p, cplx w, cplx a, cplx b, const cplx *Z, double ux,
                              double uy, int sign, cplx *acc) {
  const cplx e1 = {ux, sign*uy};
  const cplx mb = {-b.re, -b.im};
  cplx ph = {1.0, 0.0};
  for (int m = 0; m <= p + 1; ++m) {
    const cplx f = cmul(Z[m], ph);
    acc_add(p, m, cmul(w, f), acc);
    acc_add(p, m + 1, cmul(a, f), acc);
    acc_add(p, m - 1, cmul(mb, f), acc);
    if (m > 0) {
      // F_-m = (-1)^m Z_m e^{-i sg m th}
      cplx fn = cmul(Z[m], conj(ph));
      if (m & 1) fn = cscale(fn, -1.0);
      acc_add(p, -m, cmul(w, fn), acc);
      acc_add(p, -m + 1, cmul(a, fn), acc);
      acc_add(p, -m - 1, cmul(mb, fn), acc);
    }
    ph = cmul(ph, e1);
  }
}

// {{{ P2M / P2L: one wave per box, lanes over its sources

// DER: dipole weights qdx, qdy (either may be null) beside the charges
template <int KIND, bool LOCAL, bool DER>
__global__ __launch_bounds__(WAVE) void h2d_p2e_kernel(
    int p, cplx kk, int nboxes, const int *boxes, const int *starts, const int *lists,
    const int *bss, const int *bsc, const double *centers, int ncenters, const double *sx,
    const double *sy, const double *q, const double *qdx, const double *qdy, double rscale,
    double *out, int base) {
  const int ib = blockIdx.x;
  if (ib >= nboxes) return;
  const int box = boxes[ib];
  const int lane = threadIdx.x;
  const double cx = centers[box], cy = centers[ncenters + box];
  cplx acc[2*PMAX + 1];
  for (int c = 0; c < 2*p + 1; ++c) acc[c] = {0.0, 0.0};
  cplx Z[PMAX + 2];
  // P2M: the box's own sources; P2L: the sources of its list-4 boxes
  const int e0 = LOCAL ? starts[ib] : 0, e1 = LOCAL ? starts[ib + 1] : 1;
  for (int e = e0; e < e1; ++e) {
    const int sb = LOCAL ? lists[e] : box;
    const int s0 = bss[sb], ns = bsc[sb];
    for (int i = lane; i < ns; i += WAVE) {
      const int s = s0 + i;
      double ux, uy;  // e^{i th_s}: the source about the centre
      const double rho = unit(sx[s] - cx, sy[s] - cy, ux, uy);
      const cplx w = {q[2*s], q[2*s + 1]};
      if (DER) {
        const int sg = LOCAL ? 1 : -1;
        const cplx dx = qdx ? cplx{qdx[2*s], qdx[2*s + 1]} : cplx{0.0, 0.0};
        const cplx dy = qdy ? cplx{qdy[2*s], qdy[2*s + 1]} : cplx{0.0, 0.0};
        const cplx h = cscale(sarg<KIND>(kk), -0.5);
        const cplx idy = {-sg*dy.im, sg*dy.re};  // i sg dy
        const cplx a = cmul(h, cadd(dx, idy)), b = cmul(h, cplx{dx.re - idy.re, dx.im - idy.im});
        if (LOCAL) seq_h<KIND>(kk, rho, p + 1, Z);
        else seq_j<KIND>(kk, rho, p + 1, Z);
        add_terms_der(p, w, a, b, Z, ux, uy, sg, acc);
      } else if (LOCAL) {
        seq_h<KIND>(kk, rho, p > 0 ? p : 1, Z);
        add_terms(p, w, Z, ux, uy, 1, acc);
      } else {
        seq_j<KIND>(kk, rho, p, Z);
        add_terms(p, w, Z, ux, uy, -1, acc);
      }
    }
  }
  double mine_re = 0.0, mine_im = 0.0;
  for (int c = 0; c < 2*p + 1; ++c) {
    const double re = wave_sum(acc[c].re), im = wave_sum(acc[c].im);
    if (lane == c) {
      mine_re = re;
      mine_im = im;
    }
  }
  if (lane < 2*p + 1) {
    const int ac = lane > p ? lane - p : p - lane;
    const double s = LOCAL ? ipow(rscale, ac) : 1.0/ipow(rscale, ac);
    store_c(out, (long)(box - base)*(2*p + 1) + lane, {mine_re*s, mine_im*s});
  }
}

// }}}

// {{{ M2M / L2L: one wave per target box

// J_n(s|d|) and e^{i n th_d} for n = 0..N into LDS (lane 0), the same wave
template <int KIND>
__device__ void shift_data(cplx kk, double dx, double dy, int N, cplx *Js, cplx *ph) {
  if (threadIdx.x == 0) {
    double ux, uy;
    const double dl = unit(dx, dy, ux, uy);
    cplx Z[NSEQ];
    seq_j<KIND>(kk, dl, N, Z);
    cplx e = {1.0, 0.0};
    for (int n = 0; n <= N; ++n) {
      Js[n] = Z[n];
      ph[n] = e;
      e = cmul(e, cplx{ux, uy});
    }
  }
  __syncthreads();
}

// J_n e^{i s n th} for any sign of n from the n >= 0 tables (J_-n = (-1)^n J_n)
__device__ __forceinline__ cplx shift_term(int n, const cplx *Js, const cplx *ph, int sign) {
  const int an = n < 0 ? -n : n;
  cplx j = Js[an];
  if (n < 0 && (an & 1)) j = cscale(j, -1.0);
  cplx e = ph[an];
  if ((n < 0) != (sign < 0)) e = conj(e);
  return cmul(e, j);
}

template <int KIND, bool LOCAL>
__global__ __launch_bounds__(WAVE) void h2d_shift_kernel(
    int ps, int pt, cplx kk, int ntgt, const int *target_boxes, const int *relatives,
    int nboxes, const double *centers, int ncenters, double src_rscale, double tgt_rscale,
    const double *src, int sbase, double *tgt, int tbase) {
  __shared__ cplx Js[NSEQ];
  __shared__ cplx ph[NSEQ];
  __shared__ cplx S[2*PMAX + 1];
  const int it = blockIdx.x;
  if (it >= ntgt) return;
  const int tb = target_boxes[it];
  const int lane = threadIdx.x;
  const int cj = lane - pt;
  cplx acc = {0.0, 0.0};
  // M2M: the (up to 4) children; L2L: the parent
  const int nrel = LOCAL ? 1 : 4;
  for (int r = 0; r < nrel; ++r) {
    const int sb = LOCAL ? relatives[tb] : relatives[(long)r*nboxes + tb];
    if (!LOCAL && sb == 0) continue;  // same on every lane
    const double dx = centers[tb] - centers[sb], dy = centers[ncenters + tb] - centers[ncenters + sb];
    shift_data<KIND>(kk, dx, dy, ps + pt, Js, ph);
    if (lane < 2*ps + 1) {
      const int am = lane > ps ? lane - ps : ps - lane;
      const double s = LOCAL ? 1.0/ipow(src_rscale, am) : ipow(src_rscale, am);
      S[lane] = cscale(load_c(src, (long)(sb - sbase)*(2*ps + 1) + lane), s);
    }
    __syncthreads();
    if (lane < 2*pt + 1)
      for (int mi = 0; mi < 2*ps + 1; ++mi)
        acc = cfma(S[mi], shift_term(mi - ps - cj, Js, ph, LOCAL ? -1 : 1), acc);
    __syncthreads();
  }
  if (lane < 2*pt + 1) {
    const int aj = cj < 0 ? -cj : cj;
    const double s = LOCAL ? ipow(tgt_rscale, aj) : 1.0/ipow(tgt_rscale, aj);
    const long o = (long)(tb - tbase)*(2*pt + 1) + lane;
    store_c(tgt, o, cadd(load_c(tgt, o), cscale(acc, s)));
  }
}

// }}}

// {{{ M2L

// H_n(s|d|) e^{i n th_d}, n = -N..N, into D[n + N] (one thread)
template <int KIND>
__device__ void hankel_row(cplx kk, double dx, double dy, int N, cplx *D) {
  double ux, uy;
  const double dl = unit(dx, dy, ux, uy);
  cplx Z[NSEQ];
  seq_h<KIND>(kk, dl, N > 0 ? N : 1, Z);
  cplx e = {1.0, 0.0};
  for (int n = 0; n <= N; ++n) {
    const cplx h = cmul(Z[n], e);
    D[N + n] = h;
    // H_-n e^{-i n th} = (-1)^n H_n conj(e)
    const cplx hn = cmul(Z[n], conj(e));
    D[N - n] = (n & 1) ? cscale(hn, -1.0) : hn;
    e = cmul(e, cplx{ux, uy});
  }
}

template <int KIND>
__global__ void h2d_m2l_data_kernel(int N, cplx kk, int ncls, const double *vecs, int nvecs,
                                    int cls_start, double *data) {
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  if (c >= ncls) return;
  cplx D[2*NSEQ];
  hankel_row<KIND>(kk, vecs[cls_start + c], vecs[nvecs + cls_start + c], N, D);
  for (int n = 0; n < 2*N + 1; ++n) store_c(data, (long)c*(2*N + 1) + n, D[n]);
}

// one wave per target box; class data from `data` (rows by class - cls_start)
// or, when data is null, the Hankel row of each pair made in LDS
template <int KIND>
__global__ __launch_bounds__(WAVE) void h2d_m2l_kernel(
    int ps, int pt, cplx kk, int ntgt, const int *target_boxes, const int *starts,
    const int *lists, const int *classes, int cls_start, const double *data,
    const double *centers, int ncenters, double src_rscale, double tgt_rscale,
    const double *src, int sbase, double *tgt, int tbase) {
  __shared__ cplx Dl[2*NSEQ];
  __shared__ cplx S[2*PMAX + 1];
  const int it = blockIdx.x;
  if (it >= ntgt) return;
  const int tb = target_boxes[it];
  const int lane = threadIdx.x;
  const int N = ps + pt;
  cplx acc = {0.0, 0.0};
  for (int e = starts[it]; e < starts[it + 1]; ++e) {
    const int sb = lists[e];
    if (!data && lane == 0)
      hankel_row<KIND>(kk, centers[tb] - centers[sb],
                       centers[ncenters + tb] - centers[ncenters + sb], N, Dl);
    if (lane < 2*ps + 1) {
      const int am = lane > ps ? lane - ps : ps - lane;
      S[lane] = cscale(load_c(src, (long)(sb - sbase)*(2*ps + 1) + lane), ipow(src_rscale, am));
    }
    __syncthreads();
    if (lane < 2*pt + 1) {
      const long row = data ? (long)(classes[e] - cls_start)*(2*N + 1) : 0;
      for (int mi = 0; mi < 2*ps + 1; ++mi) {
        const cplx d = data ? load_c(data, row + lane + mi) : Dl[lane + mi];
        acc = cfma(d, S[mi], acc);
      }
    }
    __syncthreads();
  }
  if (lane < 2*pt + 1) {
    const int cj = lane - pt, aj = cj < 0 ? -cj : cj;
    const double s = ipow(tgt_rscale, aj)*((aj & 1) ? -1.0 : 1.0);
    store_c(tgt, (long)(tb - tbase)*(2*pt + 1) + lane, cscale(acc, s));
  }
}

// }}}

// {{{ L2P / M2P: one wave per target box, lanes over its targets

// (i/4) z
__device__ __forceinline__ cplx scale_i4(cplx z) { return {-0.25*z.im, 0.25*z.re}; }

// DER: the gradient into gx, gy and the potential into out, any of them null
template <int KIND, bool MPOLE, bool DER>
__global__ __launch_bounds__(WAVE) void h2d_e2p_kernel(
    int p, cplx kk, int ntgt, const int *target_boxes, const int *starts, const int *lists,
    const int *bts, const int *btc, const double *centers, int ncenters, const double *tx,
    const double *ty, double rscale, const double *coeffs, int base, double *out, double *gx,
    double *gy) {
  __shared__ cplx S[2*PMAX + 1];
  const int it = blockIdx.x;
  if (it >= ntgt) return;
  const int tb = target_boxes[it];
  const int lane = threadIdx.x;
  const int t0 = bts[tb], nt = btc[tb];
  const int e0 = MPOLE ? starts[it] : 0, e1 = MPOLE ? starts[it + 1] : 1;
  cplx Z[PMAX + 2];
  for (int c0 = 0; c0 < nt; c0 += WAVE) {
    const int t = t0 + c0 + lane;
    const bool active = c0 + lane < nt;
    cplx acc = {0.0, 0.0}, accx = {0.0, 0.0}, accy = {0.0, 0.0};
    for (int e = e0; e < e1; ++e) {
      const int sb = MPOLE ? lists[e] : tb;
      if (lane < 2*p + 1) {
        const int ac = lane > p ? lane - p : p - lane;
        const double s = MPOLE ? ipow(rscale, ac) : 1.0/ipow(rscale, ac);
        S[lane] = cscale(load_c(coeffs, (long)(sb - base)*(2*p + 1) + lane), s);
      }
      __syncthreads();
      if (active) {
        double ux, uy;
        const double rho = unit(tx[t] - centers[sb], ty[t] - centers[ncenters + sb], ux, uy);
        // M2P: e^{i c th}, H_c; L2P: e^{-i c th}, J_c
        const cplx e1u = {ux, MPOLE ? uy : -uy};
        if (DER) {
          // pot = sum_c S_c F_c; grad via the shifts of F (add_terms_der):
          // d_x -> (s/2) sum_m F_m (S_{m+1} - S_{m-1}),
          // d_y -> (i sg s/2) sum_m F_m (S_{m+1} + S_{m-1})
          if (MPOLE) seq_h<KIND>(kk, rho, p + 1, Z);
          else seq_j<KIND>(kk, rho, p + 1, Z);
          auto sget = [&](int c) { return (c >= -p && c <= p) ? S[p + c] : cplx{0.0, 0.0}; };
          cplx ph = {1.0, 0.0};
          for (int m = 0; m <= p + 1; ++m) {
            const cplx f = cmul(Z[m], ph);
            acc = cfma(sget(m), f, acc);
            const cplx up = sget(m + 1), dn = sget(m - 1);
            accx = cfma(f, cplx{up.re - dn.re, up.im - dn.im}, accx);
            accy = cfma(f, cadd(up, dn), accy);
            if (m > 0) {
              cplx fn = cmul(Z[m], conj(ph));
              if (m & 1) fn = cscale(fn, -1.0);
              acc = cfma(sget(-m), fn, acc);
              const cplx up2 = sget(-m + 1), dn2 = sget(-m - 1);
              accx = cfma(fn, cplx{up2.re - dn2.re, up2.im - dn2.im}, accx);
              accy = cfma(fn, cadd(up2, dn2), accy);
            }
            ph = cmul(ph, e1u);
          }
        } else {
        if (MPOLE) seq_h<KIND>(kk, rho, p > 0 ? p : 1, Z);
        else seq_j<KIND>(kk, rho, p, Z);
        cplx ph = {1.0, 0.0};
        for (int c = 0; c <= p; ++c) {
          const cplx z = Z[c];
          acc = cfma(S[p + c], cmul(z, ph), acc);
          if (c > 0) {
            const cplx tn = cmul(S[p - c], cmul(z, conj(ph)));
            acc = (c & 1) ? cplx{acc.re - tn.re, acc.im - tn.im} : cadd(acc, tn);
          }
          ph = cmul(ph, e1u);
        }
        }
      }
      __syncthreads();
    }
    if (active) {
      // M2P adds to the targets' values; L2P writes them
      auto put = [&](double *o, cplx v) {
        if (o) store_c(o, t, MPOLE ? cadd(load_c(o, t), v) : v);
      };
      put(out, scale_i4(acc));
      if (DER) {
        const cplx h = cscale(sarg<KIND>(kk), 0.5);
        const cplx ih = MPOLE ? cplx{-h.im, h.re} : cplx{h.im, -h.re};  // i sg s/2
        put(gx, scale_i4(cmul(h, accx)));
        put(gy, scale_i4(cmul(ih, accy)));
      }
    }
  }
}

// }}}

// {{{ P2P

// J0(x), Y0(x) for 0 < x <= 4 by their power series,
//   J0 = sum_m (-x^2/4)^m / (m!)^2,
//   Y0 = (2/pi) ((ln(x/2) + gamma) J0 + sum_{m>=1} (-1)^(m+1) H_m (x^2/4)^m / (m!)^2)
// with H_m the harmonic numbers; 17 terms reach 1e-19 at x = 4 (against
// scipy: 8e-16 on (0, 4]); the coefficients fold to constants when unrolled.
__device__ __forceinline__ void h0_series(double x, double &j0, double &y0) {
  const double q = 0.25*x*x;
  double t = 1.0, j = 1.0, s = 0.0, h = 0.0;
#pragma unroll
  for (int m = 1; m <= 17; ++m) {
    t *= -q*(1.0/((double)m*(double)m));
    h += 1.0/(double)m;
    j += t;
    s -= t*h;
  }
  j0 = j;
  y0 = (2.0/PI)*((log(0.5*x) + EULER_GAMMA)*j + s);
}

// (i/4) H0(k r) for r > 0: the power series up to k r = 4, the Miller /
// Neumann route above it (complex k: the complex Miller / Neumann route);
// Yukawa: K0(lam r) / (2 pi)
template <int KIND>
__device__ __forceinline__ cplx green(cplx kk, double r) {
  if (KIND == 2) {
    cplx J[2], Y[2];
    bessel_jy_c(cmul(kk, cplx{r, 0.0}), 1, J, Y);
    const cplx h = {J[0].re - Y[0].im, J[0].im + Y[0].re};  // H0 = J0 + i Y0
    return {-0.25*h.im, 0.25*h.re};
  }
  const double x = kk.re*r;
  if (KIND == 1) {
    double I[2], K[2];
    bessel_ik(x, 1, I, K);
    return {K[0]/(2.0*PI), 0.0};
  }
  double j0, y0;
  if (x <= 4.0) {
    h0_series(x, j0, y0);
  } else {
    double J[2], Y[2];
    bessel_jy(x, 1, J, Y);
    j0 = J[0];
    y0 = Y[0];
  }
  return {-0.25*y0, 0.25*j0};
}

// H0(s r), H1(s r) for r > 0 (s = k, complex k or i lam)
template <int KIND>
__device__ __forceinline__ void hankel01(cplx kk, double r, cplx &h0, cplx &h1) {
  if (KIND == 2) {
    cplx J[2], Y[2];
    bessel_jy_c(cmul(kk, cplx{r, 0.0}), 1, J, Y);
    h0 = {J[0].re - Y[0].im, J[0].im + Y[0].re};
    h1 = {J[1].re - Y[1].im, J[1].im + Y[1].re};
  } else if (KIND == 1) {
    // H_n(i x) = (2/pi) i^(-n-1) K_n(x)
    double I[2], K[2];
    bessel_ik(kk.re*r, 1, I, K);
    h0 = {0.0, -(2.0/PI)*K[0]};
    h1 = {-(2.0/PI)*K[1], 0.0};
  } else {
    double J[2], Y[2];
    bessel_jy(kk.re*r, 1, J, Y);
    h0 = {J[0], Y[0]};
    h1 = {J[1], Y[1]};
  }
}

// One pair with dipole weights u and the target gradient, d = t - src, G(r) =
// (i/4) H0(s r): the dipole term u . grad_d G = G'(r) (u . d)/r, the target
// gradient of a charge G'(r) d/r and of a dipole
// (G'' - G'/r) d (u . d)/r^2 + G' u/r, with G' = -(i/4) s H1(s r) and
// G'' - G'/r = -(i/4) s^2 H0 + (i/2) s H1/r.
template <int KIND>
__device__ __forceinline__ void pair_der(cplx kk, double dx, double dy, double r, cplx w,
                                         cplx wx, cplx wy, cplx &pot, cplx &gx, cplx &gy) {
  cplx h0, h1;
  hankel01<KIND>(kk, r, h0, h1);
  const cplx s = sarg<KIND>(kk);
  const cplx i4 = {0.0, 0.25};
  const cplx G = cmul(i4, h0);
  const cplx Gp = cscale(cmul(i4, cmul(s, h1)), -1.0);
  const cplx K2 = cadd(cscale(cmul(i4, cmul(cmul(s, s), h0)), -1.0),
                       cscale(cmul(cplx{0.0, 0.5}, cmul(s, h1)), 1.0/r));
  const double ir = 1.0/r;
  const cplx nd = cadd(cscale(wx, dx), cscale(wy, dy));  // u . d
  const cplx gpr = cscale(Gp, ir);                          // G'/r
  pot = cadd(pot, cadd(cmul(w, G), cmul(gpr, nd)));
  const cplx k2nd = cscale(cmul(K2, nd), ir*ir);
  gx = cadd(gx, cadd(cscale(cmul(w, gpr), dx), cadd(cscale(k2nd, dx), cmul(gpr, wx))));
  gy = cadd(gy, cadd(cscale(cmul(w, gpr), dy), cadd(cscale(k2nd, dy), cmul(gpr, wy))));
}

// one wave per target box, lanes over its targets; sources of list 1.  DER:
// dipole weights qdx, qdy and the gradient outputs gx, gy (any may be null)
template <int KIND, bool DER>
__global__ __launch_bounds__(WAVE) void h2d_p2p_csr_kernel(
    cplx kk, int ntgt, const int *target_boxes, const int *starts, const int *lists,
    const int *bss, const int *bsc, const int *bts, const int *btc, const double *tx,
    const double *ty, const double *sx, const double *sy, const double *q, const double *qdx,
    const double *qdy, const int *t2s, double *out, double *gx, double *gy) {
  const int it = blockIdx.x;
  if (it >= ntgt) return;
  const int tb = target_boxes[it];
  const int t0 = bts[tb], nt = btc[tb];
  for (int i = threadIdx.x; i < nt; i += WAVE) {
    const int t = t0 + i;
    const double x = tx[t], y = ty[t];
    const int self = t2s ? t2s[t] : -1;
    cplx acc = {0.0, 0.0}, ax = {0.0, 0.0}, ay = {0.0, 0.0};
    for (int e = starts[it]; e < starts[it + 1]; ++e) {
      const int sb = lists[e];
      const int s0 = bss[sb], s1 = s0 + bsc[sb];
      for (int s = s0; s < s1; ++s) {
        const double dx = x - sx[s], dy = y - sy[s];
        const double r = sqrt(dx*dx + dy*dy);
        if (r == 0.0 || s == self) continue;
        if (DER)
          pair_der<KIND>(kk, dx, dy, r, cplx{q[2*s], q[2*s + 1]},
                         qdx ? cplx{qdx[2*s], qdx[2*s + 1]} : cplx{0.0, 0.0},
                         qdy ? cplx{qdy[2*s], qdy[2*s + 1]} : cplx{0.0, 0.0}, acc, ax, ay);
        else
          acc = cfma(green<KIND>(kk, r), cplx{q[2*s], q[2*s + 1]}, acc);
      }
    }
    if (out) store_c(out, t, acc);
    if (DER && gx) store_c(gx, t, ax);
    if (DER && gy) store_c(gy, t, ay);
  }
}

template <int KIND, bool DER>
__global__ void h2d_p2p_dense_kernel(cplx kk, int nt, const double *tx, const double *ty,
                                     int ns, const double *sx, const double *sy,
                                     const double *q, const double *qdx, const double *qdy,
                                     const int *t2s, double *out, double *gx, double *gy) {
  const int t = blockIdx.x*blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const double x = tx[t], y = ty[t];
  const int self = t2s ? t2s[t] : -1;
  cplx acc = {0.0, 0.0}, ax = {0.0, 0.0}, ay = {0.0, 0.0};
  for (int s = 0; s < ns; ++s) {
    const double dx = x - sx[s], dy = y - sy[s];
    const double r = sqrt(dx*dx + dy*dy);
    if (r == 0.0 || s == self) continue;
    if (DER)
      pair_der<KIND>(kk, dx, dy, r, cplx{q[2*s], q[2*s + 1]},
                     qdx ? cplx{qdx[2*s], qdx[2*s + 1]} : cplx{0.0, 0.0},
                     qdy ? cplx{qdy[2*s], qdy[2*s + 1]} : cplx{0.0, 0.0}, acc, ax, ay);
    else
      acc = cfma(green<KIND>(kk, r), cplx{q[2*s], q[2*s + 1]}, acc);
  }
  if (out) store_c(out, t, acc);
  if (DER && gx) store_c(gx, t, ax);
  if (DER && gy) store_c(gy, t, ay);
}

// }}}

int check_orders(int ps, int pt) {
  SUMPY_REQUIRE(ps >= 0 && ps <= PMAX && pt >= 0 && pt <= PMAX,
                "h2d: orders must be in 0..%d (got %d, %d)", PMAX, ps, pt);
  return SUMPYHIP_OK;
}

int check_k(int kind, double k, double k_im) {
  SUMPY_REQUIRE(kind == 0 || kind == 1, "h2d: kind must be 0 (Helmholtz) or 1 (Yukawa)");
  SUMPY_REQUIRE(std::isfinite(k) && std::isfinite(k_im), "h2d: non-finite k");
  SUMPY_REQUIRE(kind == 0 || k_im == 0.0, "h2d: the Yukawa lambda must be real");
  SUMPY_REQUIRE(k_im == 0.0 ? k > 0.0 : (k_im > 0.0 && k >= 0.0),
                "h2d: the Helmholtz k must be > 0 or have Re k >= 0 < Im k, the Yukawa "
                "lambda > 0");
  return SUMPYHIP_OK;
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

#define H2D_CHECK(expr)              \
  do {                               \
    const int rc_ = (expr);          \
    if (rc_ != SUMPYHIP_OK) return rc_; \
  } while (0)

// KIND (0 real k, 1 Yukawa, 2 complex k) x derivative variant of a launch
#define H2D_LAUNCH(KCALL, der)                   \
  do {                                           \
    if (kind == 1) {                             \
      if (der) KCALL(1, true); else KCALL(1, false); \
    } else if (k_im != 0.0) {                    \
      if (der) KCALL(2, true); else KCALL(2, false); \
    } else {                                     \
      if (der) KCALL(0, true); else KCALL(0, false); \
    }                                            \
  } while (0)

extern "C" int sumpyhip_h2d_p2m(int kind, int order, double k, double k_im, int nboxes,
                                const int *source_boxes, const int *box_source_starts,
                                const int *box_source_counts, const double *centers,
                                int ncenters, const double *sx, const double *sy,
                                const void *weights, const void *weights_dx,
                                const void *weights_dy, double rscale, void *out, int base,
                                void *stream) {
  H2D_CHECK(check_orders(order, order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (nboxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(weights, "h2d_p2m: null charge weights");
#define P2M_CALL(KD, DR)                                                                    \
  h2d_p2e_kernel<KD, false, DR><<<nboxes, WAVE, 0, (hipStream_t)stream>>>(                 \
      order, kk, nboxes, source_boxes, nullptr, nullptr, box_source_starts,                \
      box_source_counts, centers, ncenters, sx, sy, (const double *)weights,               \
      (const double *)weights_dx, (const double *)weights_dy, rscale, (double *)out, base)
  H2D_LAUNCH(P2M_CALL, weights_dx || weights_dy);
#undef P2M_CALL
  return check_launch("h2d_p2m");
}

extern "C" int sumpyhip_h2d_p2l(int kind, int order, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *starts, const int *lists,
                                const int *box_source_starts, const int *box_source_counts,
                                const double *centers, int ncenters, const double *sx,
                                const double *sy, const void *weights, const void *weights_dx,
                                const void *weights_dy, double rscale, void *out, int base,
                                void *stream) {
  H2D_CHECK(check_orders(order, order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(weights, "h2d_p2l: null charge weights");
#define P2L_CALL(KD, DR)                                                                    \
  h2d_p2e_kernel<KD, true, DR><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(                    \
      order, kk, ntgt, target_boxes, starts, lists, box_source_starts, box_source_counts,  \
      centers, ncenters, sx, sy, (const double *)weights, (const double *)weights_dx,      \
      (const double *)weights_dy, rscale, (double *)out, base)
  H2D_LAUNCH(P2L_CALL, weights_dx || weights_dy);
#undef P2L_CALL
  return check_launch("h2d_p2l");
}

extern "C" int sumpyhip_h2d_m2m(int kind, int src_order, int tgt_order, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *box_child_ids, int nboxes,
                                const double *centers, int ncenters, double src_rscale,
                                double tgt_rscale, const void *src, int src_base, void *tgt,
                                int tgt_base, void *stream) {
  H2D_CHECK(check_orders(src_order, tgt_order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntgt == 0) return SUMPYHIP_OK;
  if (kind == 1) h2d_shift_kernel<1, false><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, box_child_ids, nboxes, centers, ncenters,
      src_rscale, tgt_rscale, (const double *)src, src_base, (double *)tgt, tgt_base);
  else if (k_im != 0.0) h2d_shift_kernel<2, false><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, box_child_ids, nboxes, centers, ncenters,
      src_rscale, tgt_rscale, (const double *)src, src_base, (double *)tgt, tgt_base);
  else h2d_shift_kernel<0, false><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, box_child_ids, nboxes, centers, ncenters,
      src_rscale, tgt_rscale, (const double *)src, src_base, (double *)tgt, tgt_base);
  return check_launch("h2d_m2m");
}

extern "C" int sumpyhip_h2d_l2l(int kind, int src_order, int tgt_order, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *box_parent_ids,
                                const double *centers, int ncenters, double src_rscale,
                                double tgt_rscale, const void *src, int src_base, void *tgt,
                                int tgt_base, void *stream) {
  H2D_CHECK(check_orders(src_order, tgt_order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntgt == 0) return SUMPYHIP_OK;
  if (kind == 1) h2d_shift_kernel<1, true><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, box_parent_ids, 0, centers, ncenters,
      src_rscale, tgt_rscale, (const double *)src, src_base, (double *)tgt, tgt_base);
  else if (k_im != 0.0) h2d_shift_kernel<2, true><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, box_parent_ids, 0, centers, ncenters,
      src_rscale, tgt_rscale, (const double *)src, src_base, (double *)tgt, tgt_base);
  else h2d_shift_kernel<0, true><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, box_parent_ids, 0, centers, ncenters,
      src_rscale, tgt_rscale, (const double *)src, src_base, (double *)tgt, tgt_base);
  return check_launch("h2d_l2l");
}

extern "C" int sumpyhip_h2d_m2l_data(int kind, int src_order, int tgt_order, double k, double k_im, int ncls,
                                     const double *vecs, int nvecs, int cls_start, void *data,
                                     void *stream) {
  H2D_CHECK(check_orders(src_order, tgt_order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ncls == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(cls_start >= 0 && cls_start + ncls <= nvecs, "h2d_m2l_data: class range");
  if (kind == 1) h2d_m2l_data_kernel<1><<<(ncls + 63)/64, 64, 0, (hipStream_t)stream>>>(
      src_order + tgt_order, kk, ncls, vecs, nvecs, cls_start, (double *)data);
  else if (k_im != 0.0) h2d_m2l_data_kernel<2><<<(ncls + 63)/64, 64, 0, (hipStream_t)stream>>>(
      src_order + tgt_order, kk, ncls, vecs, nvecs, cls_start, (double *)data);
  else h2d_m2l_data_kernel<0><<<(ncls + 63)/64, 64, 0, (hipStream_t)stream>>>(
      src_order + tgt_order, kk, ncls, vecs, nvecs, cls_start, (double *)data);
  return check_launch("h2d_m2l_data");
}

extern "C" int sumpyhip_h2d_m2l(int kind, int src_order, int tgt_order, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *starts, const int *lists,
                                const int *classes, int cls_start, const void *data,
                                const double *centers, int ncenters, double src_rscale,
                                double tgt_rscale, const void *src, int src_base, void *tgt,
                                int tgt_base, void *stream) {
  H2D_CHECK(check_orders(src_order, tgt_order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  SUMPY_REQUIRE(!data || classes, "h2d_m2l: class data without the class list");
  if (ntgt == 0) return SUMPYHIP_OK;
  if (kind == 1) h2d_m2l_kernel<1><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, starts, lists, classes, cls_start,
      (const double *)data, centers, ncenters, src_rscale, tgt_rscale, (const double *)src,
      src_base, (double *)tgt, tgt_base);
  else if (k_im != 0.0) h2d_m2l_kernel<2><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, starts, lists, classes, cls_start,
      (const double *)data, centers, ncenters, src_rscale, tgt_rscale, (const double *)src,
      src_base, (double *)tgt, tgt_base);
  else h2d_m2l_kernel<0><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(
      src_order, tgt_order, kk, ntgt, target_boxes, starts, lists, classes, cls_start,
      (const double *)data, centers, ncenters, src_rscale, tgt_rscale, (const double *)src,
      src_base, (double *)tgt, tgt_base);
  return check_launch("h2d_m2l");
}

extern "C" int sumpyhip_h2d_l2p(int kind, int order, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *box_target_starts,
                                const int *box_target_counts, const double *centers,
                                int ncenters, const double *tx, const double *ty, double rscale,
                                const void *loc, int base, void *out, void *out_dx,
                                void *out_dy, void *stream) {
  H2D_CHECK(check_orders(order, order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntgt == 0) return SUMPYHIP_OK;
#define L2P_CALL(KD, DR)                                                                    \
  h2d_e2p_kernel<KD, false, DR><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(                   \
      order, kk, ntgt, target_boxes, nullptr, nullptr, box_target_starts, box_target_counts, \
      centers, ncenters, tx, ty, rscale, (const double *)loc, base, (double *)out,         \
      (double *)out_dx, (double *)out_dy)
  H2D_LAUNCH(L2P_CALL, out_dx || out_dy || !out);
#undef L2P_CALL
  return check_launch("h2d_l2p");
}

extern "C" int sumpyhip_h2d_m2p(int kind, int order, double k, double k_im, int ntgt,
                                const int *target_boxes, const int *starts, const int *lists,
                                const int *box_target_starts, const int *box_target_counts,
                                const double *centers, int ncenters, const double *tx,
                                const double *ty, double rscale, const void *mpoles, int base,
                                void *out, void *out_dx, void *out_dy, void *stream) {
  H2D_CHECK(check_orders(order, order));
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntgt == 0) return SUMPYHIP_OK;
#define M2P_CALL(KD, DR)                                                                    \
  h2d_e2p_kernel<KD, true, DR><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(                    \
      order, kk, ntgt, target_boxes, starts, lists, box_target_starts, box_target_counts,  \
      centers, ncenters, tx, ty, rscale, (const double *)mpoles, base, (double *)out,      \
      (double *)out_dx, (double *)out_dy)
  H2D_LAUNCH(M2P_CALL, out_dx || out_dy || !out);
#undef M2P_CALL
  return check_launch("h2d_m2p");
}

extern "C" int sumpyhip_h2d_p2p_csr(int kind, double k, double k_im, int ntgt,
                                    const int *target_boxes, const int *starts, const int *lists,
                                    const int *box_source_starts, const int *box_source_counts,
                                    const int *box_target_starts, const int *box_target_counts,
                                    const double *tx, const double *ty, const double *sx,
                                    const double *sy, const void *weights,
                                    const void *weights_dx, const void *weights_dy,
                                    const int *t2s, void *out, void *out_dx, void *out_dy,
                                    void *stream) {
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntgt == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(weights, "h2d_p2p_csr: null charge weights");
#define CSR_CALL(KD, DR)                                                                    \
  h2d_p2p_csr_kernel<KD, DR><<<ntgt, WAVE, 0, (hipStream_t)stream>>>(                      \
      kk, ntgt, target_boxes, starts, lists, box_source_starts, box_source_counts,         \
      box_target_starts, box_target_counts, tx, ty, sx, sy, (const double *)weights,       \
      (const double *)weights_dx, (const double *)weights_dy, t2s, (double *)out,          \
      (double *)out_dx, (double *)out_dy)
  H2D_LAUNCH(CSR_CALL, weights_dx || weights_dy || out_dx || out_dy || !out);
#undef CSR_CALL
  return check_launch("h2d_p2p_csr");
}

extern "C" int sumpyhip_h2d_p2p_dense(int kind, double k, double k_im, int ntargets,
                                      const double *tx, const double *ty, int nsources,
                                      const double *sx, const double *sy, const void *weights,
                                      const void *weights_dx, const void *weights_dy,
                                      const int *t2s, void *out, void *out_dx, void *out_dy,
                                      void *stream) {
  H2D_CHECK(check_k(kind, k, k_im));
  const cplx kk = {k, k_im};
  if (ntargets == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(weights, "h2d_p2p_dense: null charge weights");
#define DENSE_CALL(KD, DR)                                                                  \
  h2d_p2p_dense_kernel<KD, DR><<<(ntargets + 255)/256, 256, 0, (hipStream_t)stream>>>(     \
      kk, ntargets, tx, ty, nsources, sx, sy, (const double *)weights,                     \
      (const double *)weights_dx, (const double *)weights_dy, t2s, (double *)out,          \
      (double *)out_dx, (double *)out_dy)
  H2D_LAUNCH(DENSE_CALL, weights_dx || weights_dy || out_dx || out_dy || !out);
#undef DENSE_CALL
  return check_launch("h2d_p2p_dense");
}
