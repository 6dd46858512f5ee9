// Point-to-point interactions: dense direct sum and per-box CSR near field.
//
// Reference: sumpy/p2p.py -- P2P (dense, :215-285) and P2PFromCSR (near field
// over per-box source lists, :484-833).  The reference generates one OpenCL
// kernel per kernel/derivative combination from symbolic expressions; here the
// Green's functions are written by hand and every source/target wrapper is
// folded into charge/dipole weights and potential/gradient outputs.
//
// Layout: coordinates SoA in HBM; each block stages source tiles into LDS
// (coalesced loads) and every lane owns one target (or a 1/G share of a
// target's sources when a box holds fewer than 64/G targets), so the inner
// loop reads LDS with wave-wide broadcasts.  Roofline: VALU issue (FP64 or
// FP32 pair math).  The FP32 form stores coordinates relative to the target
// box centre (or the first target of a dense block), sums each source tile in
// FP32 and folds the tile sum into a per-target FP64 accumulator.
#include "common.h"

namespace sumpyhip {
namespace {

constexpr int KLAP2 = 2, KLAP3 = 3, KHELM3 = 13;

// {{{ pair evaluators (unscaled G; d = target - source)

template <int KNL, bool DIP, bool GRAD, typename T>
__device__ __forceinline__ void pair_real(T dx, T dy, T dz, T w0, T wx, T wy, T wz,
                                          T &pot, T &g0, T &g1, T &g2) {
  if (KNL == KLAP3) {
    T r2 = dx*dx + dy*dy + dz*dz;
    T ri = rsqrt(r2);
    pot += w0*ri;
    if (DIP || GRAD) {
      T ri3 = ri*ri*ri;
      T wd = DIP ? (wx*dx + wy*dy + wz*dz) : T(0);
      if (DIP) pot -= wd*ri3;
      if (GRAD) {
        T f = -w0*ri3;
        g0 += f*dx; g1 += f*dy; g2 += f*dz;
        if (DIP) {
          T c = T(3)*wd*ri3*ri*ri;
          g0 += c*dx - wx*ri3; g1 += c*dy - wy*ri3; g2 += c*dz - wz*ri3;
        }
      }
    }
  } else {  // Laplace 2D, G = log r
    T r2 = dx*dx + dy*dy;
    pot += w0*(T(0.5)*log(r2));
    if (DIP || GRAD) {
      T inv = T(1)/r2;
      T wd = DIP ? (wx*dx + wy*dy) : T(0);
      if (DIP) pot += wd*inv;
      if (GRAD) {
        g0 += w0*dx*inv; g1 += w0*dy*inv;
        if (DIP) {
          T c = T(2)*wd*inv*inv;
          g0 += wx*inv - c*dx; g1 += wy*inv - c*dy;
        }
      }
    }
  }
}

__device__ __forceinline__ void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }

// Helmholtz 3D: G = exp(i k r)/r, complex k = kr + i ki, complex weights.
template <bool DIP, bool GRAD, typename T>
__device__ __forceinline__ void pair_helm(T dx, T dy, T dz, T kre, T kim,
                                          T w0r, T w0i, const T *wd, T *acc) {
  // acc: pot(re,im), g0(re,im), g1, g2
  T r2 = dx*dx + dy*dy + dz*dz;
  T ri = rsqrt(r2);
  T r = r2*ri;
  T s, c;
  sincos_t(kre*r, &s, &c);
  T damp = (kim != T(0)) ? exp(-kim*r) : T(1);
  T er = c*damp*ri, ei = s*damp*ri;  // e^{ikr}/r
  acc[0] += w0r*er - w0i*ei;
  acc[1] += w0r*ei + w0i*er;
  if (DIP || GRAD) {
    // f'(r)/r = e^{ikr}/r * (i k r - 1)/r^2
    T ikr_re = -kim*r, ikr_im = kre*r;  // i k r
    T ri2 = ri*ri;
    T qre = (ikr_re - T(1))*ri2, qim = ikr_im*ri2;
    T f1re = er*qre - ei*qim, f1im = er*qim + ei*qre;  // f'(r)/r
    T d[3] = {dx, dy, dz};
    if (DIP) {
      T wdr = wd[0]*dx + wd[2]*dy + wd[4]*dz, wdi = wd[1]*dx + wd[3]*dy + wd[5]*dz;
      acc[0] += wdr*f1re - wdi*f1im;
      acc[1] += wdr*f1im + wdi*f1re;
    }
    if (GRAD) {
      T gwr = w0r*f1re - w0i*f1im, gwi = w0r*f1im + w0i*f1re;
      for (int b = 0; b < 3; ++b) {
        acc[2 + 2*b] += gwr*d[b];
        acc[3 + 2*b] += gwi*d[b];
      }
      if (DIP) {
        // Hessian H_ab = f'/r (delta_ab - u_a u_b) + f'' u_a u_b, u = d/r
        // f'' = e^{ikr}/r * (-(kr)^2 - 2ikr + 2)/r^2
        T kr2re = (kre*kre - kim*kim)*r2, kr2im = T(2)*kre*kim*r2;  // (k r)^2
        T pre = (-kr2re - T(2)*ikr_re + T(2))*ri2, pim = (-kr2im - T(2)*ikr_im)*ri2;
        T f2re = er*pre - ei*pim, f2im = er*pim + ei*pre;
        T hre = f2re - f1re, him = f2im - f1im;  // coefficient of u_a u_b
        T wdr = (wd[0]*dx + wd[2]*dy + wd[4]*dz)*ri, wdi = (wd[1]*dx + wd[3]*dy + wd[5]*dz)*ri;
        for (int b = 0; b < 3; ++b) {
          T ub = d[b]*ri;
          // sum_a w_a H_ab = f1 w_b + h (w.u) u_b
          T t1r = wd[2*b]*f1re - wd[2*b+1]*f1im, t1i = wd[2*b]*f1im + wd[2*b+1]*f1re;
          T t2r = (wdr*hre - wdi*him)*ub, t2i = (wdr*him + wdi*hre)*ub;
          acc[2 + 2*b] += t1r + t2r;
          acc[3 + 2*b] += t1i + t2i;
        }
      }
    }
  }
}

// }}}

struct P2PArgs {
  double scale, kre, kim;
  int ntgt_boxes;
  const int *target_boxes, *src_box_starts, *src_box_lists;
  const int *box_src_starts, *box_src_counts, *box_tgt_starts, *box_tgt_counts;
  const double *centers;
  int nboxes;
  const double *tx, *ty, *tz, *sx, *sy, *sz;
  const double *w0, *wx, *wy, *wz;
  const int *t2s;
  double *pot, *gx, *gy, *gz;
  int ntargets, nsources;  // dense form
};

constexpr int WAVE = 64;

// {{{ CSR near field, real kernels

template <int KNL, bool DIP, bool GRAD, typename T>
__global__ __launch_bounds__(WAVE) void p2p_csr_real(P2PArgs a) {
  constexpr int DIM = KNL == KLAP3 ? 3 : 2;
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ T s_c[DIM][WAVE];
  __shared__ T s_w[DIP ? 4 : 1][WAVE];
  __shared__ int s_idx[WAVE];
  const int lane = threadIdx.x;
  for (int ib = blockIdx.x; ib < a.ntgt_boxes; ib += gridDim.x) {
    const int tb = a.target_boxes[ib];
    const int t0 = a.box_tgt_starts[tb], nt = a.box_tgt_counts[tb];
    if (nt == 0) continue;
    const int sb0 = a.src_box_starts[ib], sb1 = a.src_box_starts[ib + 1];
    int G = 1;
    while (G < WAVE && nt*(2*G) <= WAVE) G *= 2;
    const int tpc = WAVE/G;
    double org[3] = {0, 0, 0};
    if (F32) {
      // local origin: the target box centre, else its first target
      const double *tc[3] = {a.tx, a.ty, a.tz};
      for (int d = 0; d < DIM; ++d)
        org[d] = a.centers ? a.centers[(size_t)d*a.nboxes + tb] : tc[d][t0];
    }
    for (int tc = 0; tc < nt; tc += tpc) {
      const int lt = lane/G, sub = lane % G;
      const bool tvalid = tc + lt < nt;
      const int t = t0 + tc + (tvalid ? lt : 0);
      const T px = T(a.tx[t] - org[0]), py = T(a.ty[t] - org[1]),
              pz = DIM == 3 ? T(a.tz[t] - org[2]) : T(0);
      const int self = a.t2s ? a.t2s[t] : -1;
      double pot = 0, g0 = 0, g1 = 0, g2 = 0;
      for (int isb = sb0; isb < sb1; ++isb) {
        const int sbx = a.src_box_lists[isb];
        const int s0 = a.box_src_starts[sbx], ns = a.box_src_counts[sbx];
        for (int sc = 0; sc < ns; sc += WAVE) {
          __syncthreads();
          if (lane < ns - sc) {
            const int s = s0 + sc + lane;
            s_c[0][lane] = T(a.sx[s] - org[0]);
            s_c[1][lane] = T(a.sy[s] - org[1]);
            if (DIM == 3) s_c[DIM - 1][lane] = T(a.sz[s] - org[2]);
            s_w[0][lane] = T(a.w0[s]);
            if constexpr (DIP) {
              s_w[1][lane] = T(a.wx[s]);
              s_w[2][lane] = T(a.wy[s]);
              if (DIM == 3) s_w[3][lane] = T(a.wz[s]);
            }
            s_idx[lane] = s;
          }
          __syncthreads();
          const int nn = min(WAVE, ns - sc);
          T tp = 0, t0g = 0, t1g = 0, t2g = 0;
          for (int j = sub; j < nn; j += G) {
            T w0 = s_w[0][j];
            if (s_idx[j] == self) continue;
            T dx = px - s_c[0][j], dy = py - s_c[1][j];
            T dz = DIM == 3 ? pz - s_c[DIM - 1][j] : T(0);
            pair_real<KNL, DIP, GRAD, T>(dx, dy, dz, w0, DIP ? s_w[1][j] : T(0),
                                         DIP ? s_w[2][j] : T(0),
                                         (DIP && DIM == 3) ? s_w[3][j] : T(0), tp, t0g,
                                         t1g, t2g);
          }
          pot += double(tp);
          if (GRAD) { g0 += double(t0g); g1 += double(t1g); g2 += double(t2g); }
        }
      }
      pot = group_sum(pot, G);
      if (GRAD) { g0 = group_sum(g0, G); g1 = group_sum(g1, G); g2 = group_sum(g2, G); }
      if (sub == 0 && tvalid) {
        if (a.pot) a.pot[t] = a.scale*pot;
        if (GRAD) {
          a.gx[t] = a.scale*g0; a.gy[t] = a.scale*g1;
          if (DIM == 3) a.gz[t] = a.scale*g2;
        }
      }
    }
  }
}

// }}}

// {{{ CSR near field, Helmholtz (complex weights and outputs)

template <bool DIP, bool GRAD, typename T>
__global__ __launch_bounds__(WAVE) void p2p_csr_helm(P2PArgs a) {
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ T s_c[3][WAVE];
  __shared__ T s_w[DIP ? 8 : 2][WAVE];
  __shared__ int s_idx[WAVE];
  const int lane = threadIdx.x;
  const T kre = T(a.kre), kim = T(a.kim);
  for (int ib = blockIdx.x; ib < a.ntgt_boxes; ib += gridDim.x) {
    const int tb = a.target_boxes[ib];
    const int t0 = a.box_tgt_starts[tb], nt = a.box_tgt_counts[tb];
    if (nt == 0) continue;
    const int sb0 = a.src_box_starts[ib], sb1 = a.src_box_starts[ib + 1];
    int G = 1;
    while (G < WAVE && nt*(2*G) <= WAVE) G *= 2;
    const int tpc = WAVE/G;
    double org[3] = {0, 0, 0};
    if (F32) {
      const double *tc[3] = {a.tx, a.ty, a.tz};
      for (int d = 0; d < 3; ++d)
        org[d] = a.centers ? a.centers[(size_t)d*a.nboxes + tb] : tc[d][t0];
    }
    for (int tc = 0; tc < nt; tc += tpc) {
      const int lt = lane/G, sub = lane % G;
      const bool tvalid = tc + lt < nt;
      const int t = t0 + tc + (tvalid ? lt : 0);
      const T px = T(a.tx[t] - org[0]), py = T(a.ty[t] - org[1]), pz = T(a.tz[t] - org[2]);
      const int self = a.t2s ? a.t2s[t] : -1;
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int isb = sb0; isb < sb1; ++isb) {
        const int sbx = a.src_box_lists[isb];
        const int s0 = a.box_src_starts[sbx], ns = a.box_src_counts[sbx];
        for (int sc = 0; sc < ns; sc += WAVE) {
          __syncthreads();
          if (lane < ns - sc) {
            const int s = s0 + sc + lane;
            s_c[0][lane] = T(a.sx[s] - org[0]);
            s_c[1][lane] = T(a.sy[s] - org[1]);
            s_c[2][lane] = T(a.sz[s] - org[2]);
            s_w[0][lane] = T(a.w0[2*s]);
            s_w[1][lane] = T(a.w0[2*s + 1]);
            if constexpr (DIP) {
              s_w[2][lane] = T(a.wx[2*s]); s_w[3][lane] = T(a.wx[2*s + 1]);
              s_w[4][lane] = T(a.wy[2*s]); s_w[5][lane] = T(a.wy[2*s + 1]);
              s_w[6][lane] = T(a.wz[2*s]); s_w[7][lane] = T(a.wz[2*s + 1]);
            }
            s_idx[lane] = s;
          }
          __syncthreads();
          const int nn = min(WAVE, ns - sc);
          T tacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          for (int j = sub; j < nn; j += G) {
            if (s_idx[j] == self) continue;
            T wd[6];
            if (DIP) for (int q = 0; q < 6; ++q) wd[q] = s_w[2 + q][j];
            pair_helm<DIP, GRAD, T>(px - s_c[0][j], py - s_c[1][j], pz - s_c[2][j], kre,
                                    kim, s_w[0][j], s_w[1][j], wd, tacc);
          }
          for (int q = 0; q < (GRAD ? 8 : 2); ++q) acc[q] += double(tacc[q]);
        }
      }
      for (int q = 0; q < (GRAD ? 8 : 2); ++q) acc[q] = group_sum(acc[q], G);
      if (sub == 0 && tvalid) {
        if (a.pot) { a.pot[2*t] = a.scale*acc[0]; a.pot[2*t + 1] = a.scale*acc[1]; }
        if (GRAD) {
          double *g[3] = {a.gx, a.gy, a.gz};
          for (int b = 0; b < 3; ++b) {
            g[b][2*t] = a.scale*acc[2 + 2*b];
            g[b][2*t + 1] = a.scale*acc[3 + 2*b];
          }
        }
      }
    }
  }
}

// }}}

// {{{ dense direct sum: 256 targets per block, 256-source LDS tiles

constexpr int DB = 256;

template <int KNL, bool DIP, bool GRAD, typename T>
__global__ __launch_bounds__(DB) void p2p_dense_real(P2PArgs a) {
  constexpr int DIM = KNL == KLAP3 ? 3 : 2;
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ T s_c[DIM][DB];
  __shared__ T s_w[DIP ? 4 : 1][DB];
  const int tid = threadIdx.x;
  const int tblk = blockIdx.x*DB;
  const int t = min(tblk + tid, a.ntargets - 1);
  const bool tvalid = tblk + tid < a.ntargets;
  double org[3] = {0, 0, 0};
  if (F32) {
    org[0] = a.tx[tblk]; org[1] = a.ty[tblk];
    if (DIM == 3) org[2] = a.tz[tblk];
  }
  const T px = T(a.tx[t] - org[0]), py = T(a.ty[t] - org[1]),
          pz = DIM == 3 ? T(a.tz[t] - org[2]) : T(0);
  const int self = a.t2s ? a.t2s[t] : -1;
  double pot = 0, g0 = 0, g1 = 0, g2 = 0;
  for (int sc = 0; sc < a.nsources; sc += DB) {
    __syncthreads();
    if (sc + tid < a.nsources) {
      const int s = sc + tid;
      s_c[0][tid] = T(a.sx[s] - org[0]);
      s_c[1][tid] = T(a.sy[s] - org[1]);
      if (DIM == 3) s_c[DIM - 1][tid] = T(a.sz[s] - org[2]);
      s_w[0][tid] = T(a.w0[s]);
      if constexpr (DIP) {
        s_w[1][tid] = T(a.wx[s]); s_w[2][tid] = T(a.wy[s]);
        if (DIM == 3) s_w[3][tid] = T(a.wz[s]);
      }
    }
    __syncthreads();
    const int nn = min(DB, a.nsources - sc);
    T tp = 0, t0g = 0, t1g = 0, t2g = 0;
    if (self >= sc && self < sc + nn) {
      for (int j = 0; j < nn; ++j) {
        if (sc + j == self) continue;
        pair_real<KNL, DIP, GRAD, T>(px - s_c[0][j], py - s_c[1][j],
                                     DIM == 3 ? pz - s_c[DIM - 1][j] : T(0), s_w[0][j],
                                     DIP ? s_w[1][j] : T(0), DIP ? s_w[2][j] : T(0),
                                     (DIP && DIM == 3) ? s_w[3][j] : T(0), tp, t0g, t1g, t2g);
      }
    } else {
      #pragma unroll 4
      for (int j = 0; j < nn; ++j) {
        pair_real<KNL, DIP, GRAD, T>(px - s_c[0][j], py - s_c[1][j],
                                     DIM == 3 ? pz - s_c[DIM - 1][j] : T(0), s_w[0][j],
                                     DIP ? s_w[1][j] : T(0), DIP ? s_w[2][j] : T(0),
                                     (DIP && DIM == 3) ? s_w[3][j] : T(0), tp, t0g, t1g, t2g);
      }
    }
    pot += double(tp);
    if (GRAD) { g0 += double(t0g); g1 += double(t1g); g2 += double(t2g); }
  }
  if (tvalid) {
    if (a.pot) a.pot[t] = a.scale*pot;
    if (GRAD) {
      a.gx[t] = a.scale*g0; a.gy[t] = a.scale*g1;
      if (DIM == 3) a.gz[t] = a.scale*g2;
    }
  }
}

template <bool DIP, bool GRAD, typename T>
__global__ __launch_bounds__(DB) void p2p_dense_helm(P2PArgs a) {
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ T s_c[3][DB];
  __shared__ T s_w[DIP ? 8 : 2][DB];
  const int tid = threadIdx.x;
  const int tblk = blockIdx.x*DB;
  const int t = min(tblk + tid, a.ntargets - 1);
  const bool tvalid = tblk + tid < a.ntargets;
  double org[3] = {0, 0, 0};
  if (F32) { org[0] = a.tx[tblk]; org[1] = a.ty[tblk]; org[2] = a.tz[tblk]; }
  const T px = T(a.tx[t] - org[0]), py = T(a.ty[t] - org[1]), pz = T(a.tz[t] - org[2]);
  const T kre = T(a.kre), kim = T(a.kim);
  const int self = a.t2s ? a.t2s[t] : -1;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int sc = 0; sc < a.nsources; sc += DB) {
    __syncthreads();
    if (sc + tid < a.nsources) {
      const int s = sc + tid;
      s_c[0][tid] = T(a.sx[s] - org[0]);
      s_c[1][tid] = T(a.sy[s] - org[1]);
      s_c[2][tid] = T(a.sz[s] - org[2]);
      s_w[0][tid] = T(a.w0[2*s]); s_w[1][tid] = T(a.w0[2*s + 1]);
      if constexpr (DIP) {
        s_w[2][tid] = T(a.wx[2*s]); s_w[3][tid] = T(a.wx[2*s + 1]);
        s_w[4][tid] = T(a.wy[2*s]); s_w[5][tid] = T(a.wy[2*s + 1]);
        s_w[6][tid] = T(a.wz[2*s]); s_w[7][tid] = T(a.wz[2*s + 1]);
      }
    }
    __syncthreads();
    const int nn = min(DB, a.nsources - sc);
    T tacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nn; ++j) {
      if (sc + j == self) continue;
      T wd[6];
      if (DIP) for (int q = 0; q < 6; ++q) wd[q] = s_w[2 + q][j];
      pair_helm<DIP, GRAD, T>(px - s_c[0][j], py - s_c[1][j], pz - s_c[2][j], kre, kim,
                              s_w[0][j], s_w[1][j], wd, tacc);
    }
    for (int q = 0; q < (GRAD ? 8 : 2); ++q) acc[q] += double(tacc[q]);
  }
  if (tvalid) {
    if (a.pot) { a.pot[2*t] = a.scale*acc[0]; a.pot[2*t + 1] = a.scale*acc[1]; }
    if (GRAD) {
      double *g[3] = {a.gx, a.gy, a.gz};
      for (int b = 0; b < 3; ++b) {
        g[b][2*t] = a.scale*acc[2 + 2*b];
        g[b][2*t + 1] = a.scale*acc[3 + 2*b];
      }
    }
  }
}

// }}}

// {{{ dispatch

template <bool DIP, bool GRAD>
int launch_csr(int knl, int prec, const P2PArgs &a, hipStream_t st) {
  const int grid = a.ntgt_boxes < 16*num_cus() ? a.ntgt_boxes : 16*num_cus();
  if (knl == KLAP3) {
    if (prec == 0) p2p_csr_real<KLAP3, DIP, GRAD, double><<<grid, WAVE, 0, st>>>(a);
    else p2p_csr_real<KLAP3, DIP, GRAD, float><<<grid, WAVE, 0, st>>>(a);
  } else if (knl == KLAP2) {
    if (prec == 0) p2p_csr_real<KLAP2, DIP, GRAD, double><<<grid, WAVE, 0, st>>>(a);
    else p2p_csr_real<KLAP2, DIP, GRAD, float><<<grid, WAVE, 0, st>>>(a);
  } else {
    if (prec == 0) p2p_csr_helm<DIP, GRAD, double><<<grid, WAVE, 0, st>>>(a);
    else p2p_csr_helm<DIP, GRAD, float><<<grid, WAVE, 0, st>>>(a);
  }
  return check_launch("p2p_csr");
}

template <bool DIP, bool GRAD>
int launch_dense(int knl, int prec, const P2PArgs &a, hipStream_t st) {
  const int grid = (a.ntargets + DB - 1)/DB;
  if (knl == KLAP3) {
    if (prec == 0) p2p_dense_real<KLAP3, DIP, GRAD, double><<<grid, DB, 0, st>>>(a);
    else p2p_dense_real<KLAP3, DIP, GRAD, float><<<grid, DB, 0, st>>>(a);
  } else if (knl == KLAP2) {
    if (prec == 0) p2p_dense_real<KLAP2, DIP, GRAD, double><<<grid, DB, 0, st>>>(a);
    else p2p_dense_real<KLAP2, DIP, GRAD, float><<<grid, DB, 0, st>>>(a);
  } else {
    if (prec == 0) p2p_dense_helm<DIP, GRAD, double><<<grid, DB, 0, st>>>(a);
    else p2p_dense_helm<DIP, GRAD, float><<<grid, DB, 0, st>>>(a);
  }
  return check_launch("p2p_dense");
}

int check_common(int knl, int prec, const P2PArgs &a) {
  SUMPY_REQUIRE(knl == KLAP2 || knl == KLAP3 || knl == KHELM3, "unknown kernel id %d", knl);
  SUMPY_REQUIRE(prec == 0 || prec == 1, "unknown precision %d", prec);
  SUMPY_REQUIRE(a.tx && a.ty && a.sx && a.sy && a.w0, "null coordinate/weight pointer");
  if (knl != KLAP2) SUMPY_REQUIRE(a.tz && a.sz, "null z coordinate pointer");
  const bool dip = a.wx != nullptr;
  if (dip) SUMPY_REQUIRE(a.wy && (knl == KLAP2 || a.wz), "incomplete dipole weights");
  const bool grad = a.gx != nullptr;
  if (grad) SUMPY_REQUIRE(a.gy && (knl == KLAP2 || a.gz), "incomplete gradient outputs");
  SUMPY_REQUIRE(a.pot || grad, "no output requested");
  return SUMPYHIP_OK;
}

// }}}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" int sumpyhip_p2p_dense(int kernel_id, int precision, double k_re, double k_im,
                                  double scale, int ntargets, const double *tx,
                                  const double *ty, const double *tz, int nsources,
                                  const double *sx, const double *sy, const double *sz,
                                  const double *w0, const double *wx, const double *wy,
                                  const double *wz, const int *target_to_source,
                                  double *pot, double *gx, double *gy, double *gz,
                                  void *stream) {
  P2PArgs a{};
  a.scale = scale; a.kre = k_re; a.kim = k_im;
  a.ntargets = ntargets; a.nsources = nsources;
  a.tx = tx; a.ty = ty; a.tz = tz; a.sx = sx; a.sy = sy; a.sz = sz;
  a.w0 = w0; a.wx = wx; a.wy = wy; a.wz = wz; a.t2s = target_to_source;
  a.pot = pot; a.gx = gx; a.gy = gy; a.gz = gz;
  if (ntargets == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntargets > 0 && nsources >= 0, "negative sizes");
  if (nsources == 0) {
    SUMPY_REQUIRE(false, "p2p_dense: nsources must be >= 1 (reference assumes nsources>=1)");
  }
  int rc = check_common(kernel_id, precision, a);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const bool dip = wx != nullptr, grad = gx != nullptr;
  if (dip && grad) return launch_dense<true, true>(kernel_id, precision, a, st);
  if (dip) return launch_dense<true, false>(kernel_id, precision, a, st);
  if (grad) return launch_dense<false, true>(kernel_id, precision, a, st);
  return launch_dense<false, false>(kernel_id, precision, a, st);
}

extern "C" int sumpyhip_p2p_csr(int kernel_id, int precision, double k_re, double k_im,
                                double scale, int ntgt_boxes, const int *target_boxes,
                                const int *source_box_starts, const int *source_box_lists,
                                const int *box_source_starts,
                                const int *box_source_counts_nonchild,
                                const int *box_target_starts,
                                const int *box_target_counts_nonchild,
                                const double *box_centers, int nboxes, const double *tx,
                                const double *ty, const double *tz, const double *sx,
                                const double *sy, const double *sz, const double *w0,
                                const double *wx, const double *wy, const double *wz,
                                const int *target_to_source, double *pot, double *gx,
                                double *gy, double *gz, void *stream) {
  P2PArgs a{};
  a.scale = scale; a.kre = k_re; a.kim = k_im;
  a.ntgt_boxes = ntgt_boxes; a.target_boxes = target_boxes;
  a.src_box_starts = source_box_starts; a.src_box_lists = source_box_lists;
  a.box_src_starts = box_source_starts; a.box_src_counts = box_source_counts_nonchild;
  a.box_tgt_starts = box_target_starts; a.box_tgt_counts = box_target_counts_nonchild;
  a.centers = box_centers; a.nboxes = nboxes;
  a.tx = tx; a.ty = ty; a.tz = tz; a.sx = sx; a.sy = sy; a.sz = sz;
  a.w0 = w0; a.wx = wx; a.wy = wy; a.wz = wz; a.t2s = target_to_source;
  a.pot = pot; a.gx = gx; a.gy = gy; a.gz = gz;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(ntgt_boxes > 0, "negative box count");
  SUMPY_REQUIRE(target_boxes && source_box_starts && source_box_lists && box_source_starts &&
                    box_source_counts_nonchild && box_target_starts &&
                    box_target_counts_nonchild,
                "null tree pointer");
  if (box_centers) SUMPY_REQUIRE(nboxes > 0, "box_centers given without nboxes");
  int rc = check_common(kernel_id, precision, a);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const bool dip = wx != nullptr, grad = gx != nullptr;
  if (dip && grad) return launch_csr<true, true>(kernel_id, precision, a, st);
  if (dip) return launch_csr<true, false>(kernel_id, precision, a, st);
  if (grad) return launch_csr<false, true>(kernel_id, precision, a, st);
  return launch_csr<false, false>(kernel_id, precision, a, st);
}
