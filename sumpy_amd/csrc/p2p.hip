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

// precision-typed math (a plain rsqrt/log/exp call on a float argument
// resolves to the double overload).  FP32 rsqrt is the bare v_rsq_f32 (1 ulp;
// no rescaling of subnormal inputs, which a squared distance between
// distinct points of a box neighbourhood does not reach).
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_t(double x) { return rsqrt(x); }
__device__ __forceinline__ float log_t(float x) { return logf(x); }
__device__ __forceinline__ double log_t(double x) { return log(x); }
__device__ __forceinline__ float exp_t(float x) { return expf(x); }
__device__ __forceinline__ double exp_t(double x) { return exp(x); }

// {{{ pair evaluators (unscaled G; d = target - source)

template <int KNL, bool DIP, bool GRAD, typename T>
__device__ __forceinline__ void pair_real(T dx, T dy, T dz, T w0, T wx, T wy, T wz,
                                          T &pot, T &g0, T &g1, T &g2) {
  if (KNL == KLAP3) {
    T r2 = dx*dx + dy*dy + dz*dz;
    T ri = rsqrt_t(r2);
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
    pot += w0*(T(0.5)*log_t(r2));
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

// FP64 sin and cos of one argument by Cody-Waite reduction (Cody and Waite,
// "Software Manual for the Elementary Functions", 1980) with pi/2 split into
// four doubles, the first three of 26 significant bits, so that n p_i is
// exact for |n| < 2^27 (the split is pi/2 rounded in steps, computed with
// 400-bit arithmetic), then the Taylor series of sin and cos on
// [-pi/4, pi/4] (through r^17 / 17! and r^16 / 16!; the first omitted terms
// are below 1e-17 relative).  The reduction is exact to rounding for
// |x| < 2^27 pi/2 (k r up to 2.1e8, far beyond any expansion's range);
// larger arguments give NaN rather than an inaccurate value.
__device__ __forceinline__ void sincos_cw(double x, double *sn, double *cs) {
  constexpr double two_over_pi = 0.6366197723675814;
  constexpr double p1 = 0x1.921fb58000000p+0, p2 = -0x1.dde9740000000p-27,
                   p3 = 0x1.1a62630000000p-54, p4 = 0x1.8a2e03707344ap-81;
  const double n = rint(x*two_over_pi);
  double r = fma(-n, p1, x);
  r = fma(-n, p2, r);
  r = fma(-n, p3, r);
  r = fma(-n, p4, r);
  r = fabs(n) < 134217728.0 ? r : __builtin_nan("");
  const double r2 = r*r;
  // sin r = r sum_k (-1)^k r^2k / (2k+1)!, cos r = sum_k (-1)^k r^2k / (2k)!
  double ps = -1.0/355687428096000.0;           // -1/17!
  ps = fma(ps, r2, 1.0/1307674368000.0);         //  1/15!
  ps = fma(ps, r2, -1.0/6227020800.0);           // -1/13!
  ps = fma(ps, r2, 1.0/39916800.0);              //  1/11!
  ps = fma(ps, r2, -1.0/362880.0);               // -1/9!
  ps = fma(ps, r2, 1.0/5040.0);                  //  1/7!
  ps = fma(ps, r2, -1.0/120.0);                  // -1/5!
  ps = fma(ps, r2, 1.0/6.0);                     //  1/3!  (sign below)
  const double sr = fma(-r*r2, ps, r);
  double pc = 1.0/20922789888000.0;             //  1/16!
  pc = fma(pc, r2, -1.0/87178291200.0);          // -1/14!
  pc = fma(pc, r2, 1.0/479001600.0);             //  1/12!
  pc = fma(pc, r2, -1.0/3628800.0);              // -1/10!
  pc = fma(pc, r2, 1.0/40320.0);                 //  1/8!
  pc = fma(pc, r2, -1.0/720.0);                  // -1/6!
  pc = fma(pc, r2, 1.0/24.0);                    //  1/4!
  pc = fma(pc, r2, -0.5);                        // -1/2!
  const double cr = fma(pc, r2, 1.0);
  const int q = int(n) & 3;  // |n| < 2^27 where the result is finite
  const double s1 = (q & 1) ? cr : sr, c1 = (q & 1) ? sr : cr;
  *sn = (q & 2) ? -s1 : s1;
  *cs = ((q + 1) & 2) ? -c1 : c1;
}

// FP32 form of the same: pi/2 in four floats, the first three of 8
// significant bits (n p_i exact for |n| < 2^16, |x| < 1e5; larger arguments
// give NaN), Taylor series through r^9 / 9! and r^10 / 10! (first omitted
// terms below 2e-9).
__device__ __forceinline__ void sincos_cw(float x, float *sn, float *cs) {
  constexpr float two_over_pi = 0.63661975f;
  constexpr float p1 = 0x1.92p+0f, p2 = 0x1.fcp-12f, p3 = -0x1.58p-21f, p4 = 0x1.10b462p-30f;
  const float n = rintf(x*two_over_pi);
  float r = fmaf(-n, p1, x);
  r = fmaf(-n, p2, r);
  r = fmaf(-n, p3, r);
  r = fmaf(-n, p4, r);
  r = fabsf(n) < 65536.0f ? r : __builtin_nanf("");
  const float r2 = r*r;
  float ps = 1.0f/362880.0f;                    //  1/9!
  ps = fmaf(ps, r2, -1.0f/5040.0f);              // -1/7!
  ps = fmaf(ps, r2, 1.0f/120.0f);                //  1/5!
  ps = fmaf(ps, r2, -1.0f/6.0f);                 // -1/3!
  const float sr = fmaf(r*r2, ps, r);
  float pc = -1.0f/3628800.0f;                  // -1/10!
  pc = fmaf(pc, r2, 1.0f/40320.0f);              //  1/8!
  pc = fmaf(pc, r2, -1.0f/720.0f);               // -1/6!
  pc = fmaf(pc, r2, 1.0f/24.0f);                 //  1/4!
  pc = fmaf(pc, r2, -0.5f);                      // -1/2!
  const float cr = fmaf(pc, r2, 1.0f);
  const int q = int(n) & 3;
  const float s1 = (q & 1) ? cr : sr, c1 = (q & 1) ? sr : cr;
  *sn = (q & 2) ? -s1 : s1;
  *cs = ((q + 1) & 2) ? -c1 : c1;
}

__device__ __forceinline__ void sincos_t(double x, double *s, double *c) { sincos_cw(x, s, c); }
// FP32: the hardware sine and cosine (v_sin_f32 / v_cos_f32, argument in
// revolutions) after reduction to [-1/2, 1/2] revolutions; measured on gfx950
// within 1.3e-7 absolute of sin / cos over that interval.  (The reduction
// rounds x / 2 pi once in FP32, like the FP32 product k r itself.)
__device__ __forceinline__ void sincos_t(float x, float *s, float *c) {
  float rev = x*0.15915494309189535f;
  rev -= rintf(rev);
  *s = __builtin_amdgcn_sinf(rev);
  *c = __builtin_amdgcn_cosf(rev);
}

// Helmholtz 3D: G = exp(i k r)/r, complex k = kr + i ki (CK; real k
// otherwise: no damping factor), complex weights.
// skip: the exclude-self pair, masked through 1/r (every term carries it)
template <bool DIP, bool GRAD, typename T, bool CK = true>
__device__ __forceinline__ void pair_helm(T dx, T dy, T dz, T kre, T kim,
                                          T w0r, T w0i, const T *wd, T *acc,
                                          bool skip = false) {
  // acc: pot(re,im), g0(re,im), g1, g2
  T r2 = dx*dx + dy*dy + dz*dz;
  T ri = rsqrt_t(r2);
  ri = skip ? T(0) : ri;
  T r = r2*ri;
  T s, c;
  sincos_t(kre*r, &s, &c);
  T damp = T(1);
  if constexpr (CK) damp = exp_t(-kim*r);
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

// {{{ CSR near field, real kernels: one 256-thread block per target box

constexpr int NB = 256;     // threads per block
constexpr int MAXB = 256;   // source boxes staged per pass

// Sources per LDS chunk: (x, y, z, w0) [+ (wx, wy, wz, 0)] vectors, <= 48 KB.
// Charges: 1152 in FP32 (1024 + the run padding; 18 KB), 640 in FP64 (20 KB),
// about the neighbourhood of a 32-64-particle leaf, so that several blocks of
// a CU hide each other's list and staging latency.
template <typename T, bool DIP>
struct Chunk {
  static constexpr int N = DIP ? 48*1024/(int)(sizeof(T)*8) : (sizeof(T) == 4 ? 1152 : 640);
};

// Pair value with the exclude-self pair masked (skip): every term carries a
// factor ri (3D) or inv/lg (2D) that the mask sets to zero.
template <int KNL, bool DIP, bool GRAD, typename T>
__device__ __forceinline__ void pair_masked(T dx, T dy, T dz, T w0, T wx, T wy, T wz, bool skip,
                                            T &pot, T &g0, T &g1, T &g2) {
  if (KNL == KLAP3) {
    const T r2 = dx*dx + dy*dy + dz*dz;
    T ri = rsqrt_t(r2);
    ri = skip ? T(0) : ri;
    pot += w0*ri;
    if (DIP || GRAD) {
      const T ri3 = ri*ri*ri;
      const T wd = DIP ? (wx*dx + wy*dy + wz*dz) : T(0);
      if (DIP) pot -= wd*ri3;
      if (GRAD) {
        const T f = -w0*ri3;
        g0 += f*dx; g1 += f*dy; g2 += f*dz;
        if (DIP) {
          const T c = T(3)*wd*ri3*ri*ri;
          g0 += c*dx - wx*ri3; g1 += c*dy - wy*ri3; g2 += c*dz - wz*ri3;
        }
      }
    }
  } else {  // Laplace 2D, G = log r
    const T r2 = dx*dx + dy*dy;
    T lg = T(0.5)*log_t(r2);
    lg = skip ? T(0) : lg;
    pot += w0*lg;
    if (DIP || GRAD) {
      T inv = T(1)/r2;
      inv = skip ? T(0) : inv;
      const T wd = DIP ? (wx*dx + wy*dy) : T(0);
      if (DIP) pot += wd*inv;
      if (GRAD) {
        g0 += w0*dx*inv; g1 += w0*dy*inv;
        if (DIP) {
          const T c = T(2)*wd*inv*inv;
          g0 += wx*inv - c*dx; g1 += wy*inv - c*dy;
        }
      }
    }
  }
}

// Per target box: the source boxes of its list are staged MAXB at a time
// (starts, counts, exclusive offsets); their particles, concatenated, are
// loaded into LDS in chunks of Chunk::N (each thread finds its box by binary
// search over the offsets, so all loads of a chunk are independent).  G lanes
// share a target (G = the largest power of two <= 64 with nt*G <= 256) and
// take every G-th source; their sums are combined with cross-lane adds.  The
// exclude-self pair is found by its position in the concatenation and masked
// with a select.  FP32: coordinates relative to the box centre (or the first
// target), FP32 sums of 16 pairs folded into FP64 accumulators.
template <int KNL, bool DIP, bool GRAD, bool SELF, typename T>
__global__ __launch_bounds__(NB) void p2p_csr_real(P2PArgs a) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  constexpr int DIM = KNL == KLAP3 ? 3 : 2;
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int CH = Chunk<T, DIP>::N;
  // FP32: chunks hold CH - 128 sources so that the G runs of odd length L
  // (G L < nn + 2 G <= nn + 128) fit, padded with zero-weight far sources
  constexpr int CHN = F32 ? CH - 128 : CH;
  constexpr int FOLD = F32 ? 16 : (1 << 20);
  __shared__ v4 s_src[CH];
  __shared__ v4 s_dip[DIP ? CH : 1];
  __shared__ int s_bs[MAXB], s_bn[MAXB], s_boff[MAXB + 1];
  __shared__ int s_ktb;  // position of the target box in the source-box chunk, or -1
  __shared__ double s_red[GRAD ? 4 : 1][NB];
  const int tid = threadIdx.x;
  for (int ib = blockIdx.x; ib < a.ntgt_boxes; ib += gridDim.x) {
    const int tb = a.target_boxes[ib];
    const int t0 = a.box_tgt_starts[tb], nt = a.box_tgt_counts[tb];
    if (nt == 0) continue;
    const int sb0 = a.src_box_starts[ib], sb1 = a.src_box_starts[ib + 1];
    // G consecutive threads per target (any G, so that nt G fills the block:
    // the lane sums are added through LDS)
    const int G = max(1, min(64, NB/nt));
    const int tpp = NB/G;
    double org[3] = {0, 0, 0};
    if (F32) {
      const double *tcrd[3] = {a.tx, a.ty, a.tz};
      for (int d = 0; d < DIM; ++d)
        org[d] = a.centers ? a.centers[(size_t)d*a.nboxes + tb] : tcrd[d][t0];
    }
    for (int tc = 0; tc < nt; tc += tpp) {
      const int lt = tid/G, sub = tid % G;
      const bool tvalid = tc + lt < nt;
      const int t = t0 + tc + (tvalid ? lt : 0);
      const T px = T(a.tx[t] - org[0]), py = T(a.ty[t] - org[1]),
              pz = DIM == 3 ? T(a.tz[t] - org[2]) : T(0);
      const int self = SELF ? a.t2s[t] : -1;
      double pot = 0, g0 = 0, g1 = 0, g2 = 0;
      for (int bb = sb0; bb < sb1; bb += MAXB) {
        const int nbx = min(MAXB, sb1 - bb);
        __syncthreads();
        if (tid == 0) s_ktb = -1;
        __syncthreads();
        if (tid < nbx) {
          const int sbx = a.src_box_lists[bb + tid];
          s_bs[tid] = a.box_src_starts[sbx];
          s_bn[tid] = a.box_src_counts[sbx];
          if (SELF && sbx == tb) s_ktb = tid;
        }
        __syncthreads();
        if (tid < 64) {  // exclusive scan of the counts (4 per lane)
          int v[4], loc = 0;
          for (int i = 0; i < 4; ++i) {
            const int k = 4*tid + i;
            v[i] = k < nbx ? s_bn[k] : 0;
            loc += v[i];
          }
          int incl = loc;
          for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (tid >= o) incl += y;
          }
          int run = incl - loc;
          for (int i = 0; i < 4; ++i) { s_boff[4*tid + i] = run; run += v[i]; }
          if (tid == 63) s_boff[MAXB] = run;
        }
        __syncthreads();
        const int total = s_boff[nbx];
        int pself = -1;
        if (SELF) {
          // a target's own source usually lies in its own box; else search
          const int kt = s_ktb;
          if (kt >= 0 && (unsigned)(self - s_bs[kt]) < (unsigned)s_bn[kt]) {
            pself = s_boff[kt] + self - s_bs[kt];
          } else {
            for (int k = 0; k < nbx; ++k)
              if ((unsigned)(self - s_bs[k]) < (unsigned)s_bn[k]) pself = s_boff[k] + self - s_bs[k];
          }
        }
        for (int c0 = 0; c0 < total; c0 += CHN) {
          const int nn = min(CHN, total - c0);
          // FP32: lane sub of a target takes the contiguous run [sub L,
          // sub L + L) of the chunk, L = ceil(nn / G) rounded up to odd (the
          // G runs of a wave then start on different LDS banks)
          const int L = ((nn + G - 1)/G) | 1;
          if (c0 > 0) __syncthreads();
          if constexpr (F32) {
            // the run tails past nn: sources 1e18 away with zero weights
            // (every term of the pair underflows to or multiplies zero)
            for (int p = nn + tid; p < G*L; p += NB) {
              s_src[p] = v4{T(1e18), T(1e18), T(1e18), T(0)};
              if constexpr (DIP) s_dip[p] = v4{0, 0, 0, 0};
            }
          }
          for (int p = tid; p < nn; p += NB) {
            const int q = c0 + p;
            int lo = 0, hi = nbx;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (s_boff[mid] <= q) lo = mid; else hi = mid;
            }
            const int s = s_bs[lo] + q - s_boff[lo];
            s_src[p] = v4{T(a.sx[s] - org[0]), T(a.sy[s] - org[1]),
                          DIM == 3 ? T(a.sz[s] - org[2]) : T(0), T(a.w0[s])};
            if constexpr (DIP)
              s_dip[p] = v4{T(a.wx[s]), T(a.wy[s]), DIM == 3 ? T(a.wz[s]) : T(0), T(0)};
          }
          __syncthreads();
          const int jself = pself - c0;
          if constexpr (F32) {
            // every lane runs L pairs (wave-uniform loop control), the
            // folds of 16 unrolled at constant LDS offsets; the excluded
            // pair is masked with a select only in the folds where some
            // lane of the wave holds it
            const int iself = (SELF && jself >= 0 && jself < nn && jself >= sub*L &&
                               jself < sub*L + L) ? jself - sub*L : -1;
            const v4 *bs = s_src + sub*L;
            const v4 *bd = s_dip + (DIP ? sub*L : 0);
            auto pair = [&](int i, bool skip, T &tp, T &t0g, T &t1g, T &t2g) {
              const v4 sv = bs[i];
              v4 dv = v4{0, 0, 0, 0};
              if constexpr (DIP) dv = bd[i];
              pair_masked<KNL, DIP, GRAD, T>(px - sv.x, py - sv.y, DIM == 3 ? pz - sv.z : T(0),
                                             sv.w, dv.x, dv.y, dv.z, skip, tp, t0g, t1g, t2g);
            };
            for (int i0 = 0; i0 < L; i0 += FOLD) {
              T tp = 0, t0g = 0, t1g = 0, t2g = 0;
              if (i0 + FOLD <= L) {
                if (SELF && __any(iself >= i0 && iself < i0 + FOLD)) {
                  #pragma unroll
                  for (int k = 0; k < FOLD; ++k) pair(i0 + k, i0 + k == iself, tp, t0g, t1g, t2g);
                } else {
                  #pragma unroll
                  for (int k = 0; k < FOLD; ++k) pair(i0 + k, false, tp, t0g, t1g, t2g);
                }
              } else {
                for (int i = i0; i < L; ++i) pair(i, SELF && i == iself, tp, t0g, t1g, t2g);
              }
              pot += double(tp);
              if (GRAD) { g0 += double(t0g); g1 += double(t1g); g2 += double(t2g); }
            }
          } else
          for (int j0 = sub; j0 < nn; j0 += G*FOLD) {
            const int jend = min(nn, j0 + G*FOLD);
            T tp = 0, t0g = 0, t1g = 0, t2g = 0;
            {
              #pragma unroll 4
              for (int j = j0; j < jend; j += G) {
                const v4 sv = s_src[j];
                v4 dv = v4{0, 0, 0, 0};
                if constexpr (DIP) dv = s_dip[j];
                pair_masked<KNL, DIP, GRAD, T>(px - sv.x, py - sv.y,
                                               DIM == 3 ? pz - sv.z : T(0), sv.w, dv.x, dv.y,
                                               dv.z, SELF && j == jself, tp, t0g, t1g, t2g);
              }
            }
            pot += double(tp);
            if (GRAD) { g0 += double(t0g); g1 += double(t1g); g2 += double(t2g); }
          }
        }
      }
      __syncthreads();  // the previous pass's sums have been read
      s_red[0][tid] = pot;
      if constexpr (GRAD) { s_red[1][tid] = g0; s_red[2][tid] = g1; s_red[3][tid] = g2; }
      __syncthreads();
      if (sub == 0 && tvalid) {
        const int q0 = lt*G;
        pot = 0; g0 = g1 = g2 = 0;
        for (int q = q0; q < q0 + G; ++q) {
          pot += s_red[0][q];
          if constexpr (GRAD) { g0 += s_red[1][q]; g1 += s_red[2][q]; g2 += s_red[3][q]; }
        }
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

// Sources per LDS chunk: (x, y, z, Re w0) and Im w0 [+ the dipole weights],
// about 20 KB, as for the real kernels.
template <typename T, bool DIP>
struct HChunk {
  static constexpr int N = DIP ? (sizeof(T) == 4 ? 512 : 256) : (sizeof(T) == 4 ? 1024 : 512);
};

// The structure of p2p_csr_real: one 256-thread block per target box, its
// list's source boxes staged MAXB at a time (starts, counts, offsets), their
// particles loaded into LDS in chunks (binary search of the box: independent
// loads), G = floor(256 / nt) lanes per target taking every G-th source,
// lane sums added through LDS; the exclude-self pair masked with a select at
// its position in the concatenation.  FP32: coordinates relative to the box
// centre, FP32 sums of 16 pairs folded into FP64.
template <bool DIP, bool GRAD, bool SELF, typename T, bool CK>
__global__ __launch_bounds__(NB) void p2p_csr_helm(P2PArgs a) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  typedef T v2 __attribute__((ext_vector_type(2)));
  constexpr int CH = HChunk<T, DIP>::N;
  constexpr int NACC = GRAD ? 8 : 2;
  constexpr int FOLD = sizeof(T) == 4 ? 16 : (1 << 20);
  __shared__ v4 s_src[CH];                 // x, y, z, Re w0
  __shared__ T s_wi[CH];                   // Im w0
  __shared__ v4 s_dxy[DIP ? CH : 1];       // Re/Im wx, Re/Im wy
  __shared__ v2 s_dz[DIP ? CH : 1];        // Re/Im wz
  __shared__ int s_bs[MAXB], s_bn[MAXB], s_boff[MAXB + 1];
  __shared__ int s_ktb;
  __shared__ double s_red[NACC][NB];
  const int tid = threadIdx.x;
  const T kre = T(a.kre), kim = T(a.kim);
  for (int ib = blockIdx.x; ib < a.ntgt_boxes; ib += gridDim.x) {
    const int tb = a.target_boxes[ib];
    const int t0 = a.box_tgt_starts[tb], nt = a.box_tgt_counts[tb];
    if (nt == 0) continue;
    const int sb0 = a.src_box_starts[ib], sb1 = a.src_box_starts[ib + 1];
    const int G = max(1, min(64, NB/nt));
    const int tpp = NB/G;
    double org[3] = {0, 0, 0};
    if (sizeof(T) == 4) {
      const double *tcrd[3] = {a.tx, a.ty, a.tz};
      for (int d = 0; d < 3; ++d)
        org[d] = a.centers ? a.centers[(size_t)d*a.nboxes + tb] : tcrd[d][t0];
    }
    for (int tc = 0; tc < nt; tc += tpp) {
      const int lt = tid/G, sub = tid % G;
      const bool tvalid = tc + lt < nt;
      const int t = t0 + tc + (tvalid ? lt : 0);
      const T px = T(a.tx[t] - org[0]), py = T(a.ty[t] - org[1]), pz = T(a.tz[t] - org[2]);
      const int self = SELF ? a.t2s[t] : -1;
      double acc[NACC];
      for (int q = 0; q < NACC; ++q) acc[q] = 0;
      for (int bb = sb0; bb < sb1; bb += MAXB) {
        const int nbx = min(MAXB, sb1 - bb);
        __syncthreads();
        if (tid == 0) s_ktb = -1;
        __syncthreads();
        if (tid < nbx) {
          const int sbx = a.src_box_lists[bb + tid];
          s_bs[tid] = a.box_src_starts[sbx];
          s_bn[tid] = a.box_src_counts[sbx];
          if (SELF && sbx == tb) s_ktb = tid;
        }
        __syncthreads();
        if (tid < 64) {  // exclusive scan of the counts (4 per lane)
          int v[4], loc = 0;
          for (int i = 0; i < 4; ++i) {
            const int k = 4*tid + i;
            v[i] = k < nbx ? s_bn[k] : 0;
            loc += v[i];
          }
          int incl = loc;
          for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (tid >= o) incl += y;
          }
          int run = incl - loc;
          for (int i = 0; i < 4; ++i) { s_boff[4*tid + i] = run; run += v[i]; }
          if (tid == 63) s_boff[MAXB] = run;
        }
        __syncthreads();
        const int total = s_boff[nbx];
        int pself = -1;
        if (SELF) {
          const int kt = s_ktb;
          if (kt >= 0 && (unsigned)(self - s_bs[kt]) < (unsigned)s_bn[kt]) {
            pself = s_boff[kt] + self - s_bs[kt];
          } else {
            for (int k = 0; k < nbx; ++k)
              if ((unsigned)(self - s_bs[k]) < (unsigned)s_bn[k]) pself = s_boff[k] + self - s_bs[k];
          }
        }
        for (int c0 = 0; c0 < total; c0 += CH) {
          const int nn = min(CH, total - c0);
          if (c0 > 0) __syncthreads();
          for (int p = tid; p < nn; p += NB) {
            const int q = c0 + p;
            int lo = 0, hi = nbx;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (s_boff[mid] <= q) lo = mid; else hi = mid;
            }
            const int s = s_bs[lo] + q - s_boff[lo];
            s_src[p] = v4{T(a.sx[s] - org[0]), T(a.sy[s] - org[1]), T(a.sz[s] - org[2]),
                          T(a.w0[2*s])};
            s_wi[p] = T(a.w0[2*s + 1]);
            if constexpr (DIP) {
              s_dxy[p] = v4{T(a.wx[2*s]), T(a.wx[2*s + 1]), T(a.wy[2*s]), T(a.wy[2*s + 1])};
              s_dz[p] = v2{T(a.wz[2*s]), T(a.wz[2*s + 1])};
            }
          }
          __syncthreads();
          const int jself = pself - c0;
          for (int j0 = sub; j0 < nn; j0 += G*FOLD) {
            const int jend = min(nn, j0 + G*FOLD);
            T tacc[NACC];
            for (int q = 0; q < NACC; ++q) tacc[q] = 0;
            for (int j = j0; j < jend; j += G) {
              const v4 sv = s_src[j];
              T wd[6] = {0, 0, 0, 0, 0, 0};
              if constexpr (DIP) {
                const v4 d1 = s_dxy[j];
                const v2 d2 = s_dz[j];
                wd[0] = d1.x; wd[1] = d1.y; wd[2] = d1.z; wd[3] = d1.w; wd[4] = d2.x; wd[5] = d2.y;
              }
              pair_helm<DIP, GRAD, T, CK>(px - sv.x, py - sv.y, pz - sv.z, kre, kim, sv.w,
                                          s_wi[j], wd, tacc, SELF && j == jself);
            }
            for (int q = 0; q < NACC; ++q) acc[q] += double(tacc[q]);
          }
        }
      }
      __syncthreads();  // the previous pass's sums have been read
      for (int q = 0; q < NACC; ++q) s_red[q][tid] = acc[q];
      __syncthreads();
      if (sub == 0 && tvalid) {
        const int q0 = lt*G;
        for (int q = 0; q < NACC; ++q) {
          double v = 0;
          for (int l = q0; l < q0 + G; ++l) v += s_red[q][l];
          acc[q] = v;
        }
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

template <bool DIP, bool GRAD, typename T, bool CK>
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
    const int jself = self - sc;  // the excluded pair of this tile, masked
    for (int j = 0; j < nn; ++j) {
      T wd[6] = {0, 0, 0, 0, 0, 0};
      if (DIP) for (int q = 0; q < 6; ++q) wd[q] = s_w[2 + q][j];
      pair_helm<DIP, GRAD, T, CK>(px - s_c[0][j], py - s_c[1][j], pz - s_c[2][j], kre, kim,
                              s_w[0][j], s_w[1][j], wd, tacc, j == jself);
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

template <int KNL, bool DIP, bool GRAD, bool SELF>
void launch_csr_real(int prec, const P2PArgs &a, hipStream_t st) {
  const int grid = a.ntgt_boxes;
  if (prec == 0) p2p_csr_real<KNL, DIP, GRAD, SELF, double><<<grid, NB, 0, st>>>(a);
  else p2p_csr_real<KNL, DIP, GRAD, SELF, float><<<grid, NB, 0, st>>>(a);
}

template <bool DIP, bool GRAD, bool CK>
void launch_csr_helm(int prec, bool self, const P2PArgs &a, hipStream_t st) {
  const int hgrid = a.ntgt_boxes;  // one box per block (as the real kernels)
  if (self) {
    if (prec == 0) p2p_csr_helm<DIP, GRAD, true, double, CK><<<hgrid, NB, 0, st>>>(a);
    else p2p_csr_helm<DIP, GRAD, true, float, CK><<<hgrid, NB, 0, st>>>(a);
  } else {
    if (prec == 0) p2p_csr_helm<DIP, GRAD, false, double, CK><<<hgrid, NB, 0, st>>>(a);
    else p2p_csr_helm<DIP, GRAD, false, float, CK><<<hgrid, NB, 0, st>>>(a);
  }
}

template <bool DIP, bool GRAD>
int launch_csr(int knl, int prec, const P2PArgs &a, hipStream_t st) {
  const int grid = a.ntgt_boxes < 16*num_cus() ? a.ntgt_boxes : 16*num_cus();
  const bool self = a.t2s != nullptr;
  if (knl == KLAP3) {
    if (self) launch_csr_real<KLAP3, DIP, GRAD, true>(prec, a, st);
    else launch_csr_real<KLAP3, DIP, GRAD, false>(prec, a, st);
  } else if (knl == KLAP2) {
    if (self) launch_csr_real<KLAP2, DIP, GRAD, true>(prec, a, st);
    else launch_csr_real<KLAP2, DIP, GRAD, false>(prec, a, st);
  } else {
    if (a.kim != 0.0) launch_csr_helm<DIP, GRAD, true>(prec, self, a, st);
    else launch_csr_helm<DIP, GRAD, false>(prec, self, a, st);
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
    if (a.kim != 0.0) {
      if (prec == 0) p2p_dense_helm<DIP, GRAD, double, true><<<grid, DB, 0, st>>>(a);
      else p2p_dense_helm<DIP, GRAD, float, true><<<grid, DB, 0, st>>>(a);
    } else {
      if (prec == 0) p2p_dense_helm<DIP, GRAD, double, false><<<grid, DB, 0, st>>>(a);
      else p2p_dense_helm<DIP, GRAD, float, false><<<grid, DB, 0, st>>>(a);
    }
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

// {{{ interaction matrices: one entry per thread

// Entry e is the pair (itgt, isrc) = (tidx[e], sidx[e]), or (e / ns, e % ns)
// for the full row-major matrix (tidx null).  The entry holds the unscaled
// kernel of the pair with the source's unit-strength weights (charge w0 and
// dipole wd, the source kernels folded as in the direct sums) times scale:
// its potential and, with GRAD, its target gradient.  A self pair (exclude
// self, t2s[itgt] == isrc) is zero.  Reference: P2PMatrixGenerator and
// P2PMatrixSubsetGenerator (sumpy/p2p.py:292-477).
struct MatArgs {
  double scale, kre, kim;
  int nt, ns;
  long long nentries;
  const double *tx, *ty, *tz, *sx, *sy, *sz;
  const double *w0, *wx, *wy, *wz;
  const int *t2s, *tidx, *sidx;
  double *pot, *gx, *gy, *gz;
};

__device__ __forceinline__ void entry_pair(const MatArgs &a, long long e, int &it, int &is) {
  if (a.tidx) {
    it = a.tidx[e];
    is = a.sidx[e];
  } else {
    it = (int)(e/a.ns);
    is = (int)(e - (long long)it*a.ns);
  }
}

template <int KNL, bool DIP, bool GRAD>
__global__ __launch_bounds__(256) void p2p_matrix_real(MatArgs a) {
  for (long long e = blockIdx.x*256ll + threadIdx.x; e < a.nentries;
       e += (long long)gridDim.x*256) {
    int it, is;
    entry_pair(a, e, it, is);
    double pot = 0, g0 = 0, g1 = 0, g2 = 0;
    if (!(a.t2s && a.t2s[it] == is)) {
      const double dx = a.tx[it] - a.sx[is], dy = a.ty[it] - a.sy[is];
      const double dz = KNL == KLAP3 ? a.tz[it] - a.sz[is] : 0.0;
      pair_real<KNL, DIP, GRAD, double>(dx, dy, dz, a.w0[is], DIP ? a.wx[is] : 0.0,
                                        DIP ? a.wy[is] : 0.0,
                                        DIP && KNL == KLAP3 ? a.wz[is] : 0.0, pot, g0, g1, g2);
    }
    if (a.pot) a.pot[e] = a.scale*pot;
    if (GRAD) {
      a.gx[e] = a.scale*g0;
      a.gy[e] = a.scale*g1;
      if (KNL == KLAP3) a.gz[e] = a.scale*g2;
    }
  }
}

template <bool DIP, bool GRAD>
__global__ __launch_bounds__(256) void p2p_matrix_helm(MatArgs a) {
  for (long long e = blockIdx.x*256ll + threadIdx.x; e < a.nentries;
       e += (long long)gridDim.x*256) {
    int it, is;
    entry_pair(a, e, it, is);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!(a.t2s && a.t2s[it] == is)) {
      double wd[6] = {0, 0, 0, 0, 0, 0};
      if (DIP) {
        wd[0] = a.wx[2*(size_t)is]; wd[1] = a.wx[2*(size_t)is + 1];
        wd[2] = a.wy[2*(size_t)is]; wd[3] = a.wy[2*(size_t)is + 1];
        wd[4] = a.wz[2*(size_t)is]; wd[5] = a.wz[2*(size_t)is + 1];
      }
      pair_helm<DIP, GRAD, double>(a.tx[it] - a.sx[is], a.ty[it] - a.sy[is],
                                   a.tz[it] - a.sz[is], a.kre, a.kim, a.w0[2*(size_t)is],
                                   a.w0[2*(size_t)is + 1], wd, acc);
    }
    if (a.pot) {
      a.pot[2*e] = a.scale*acc[0];
      a.pot[2*e + 1] = a.scale*acc[1];
    }
    if (GRAD) {
      double *g[3] = {a.gx, a.gy, a.gz};
      for (int b = 0; b < 3; ++b) {
        g[b][2*e] = a.scale*acc[2 + 2*b];
        g[b][2*e + 1] = a.scale*acc[3 + 2*b];
      }
    }
  }
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

extern "C" int sumpyhip_p2p_matrix(int kernel_id, double k_re, double k_im, double scale,
                                   int ntargets, const double *tx, const double *ty,
                                   const double *tz, int nsources, const double *sx,
                                   const double *sy, const double *sz, const void *w0,
                                   const void *wx, const void *wy, const void *wz,
                                   const int *target_to_source, long long nentries,
                                   const int *tgtindices, const int *srcindices, void *pot,
                                   void *gx, void *gy, void *gz, void *stream) {
  SUMPY_REQUIRE(kernel_id == 2 || kernel_id == 3 || kernel_id == 13,
                "p2p_matrix: unknown kernel %d", kernel_id);
  const int dim = kernel_id == 2 ? 2 : 3;
  SUMPY_REQUIRE(tx && ty && sx && sy && w0 && (dim == 2 || (tz && sz)),
                "p2p_matrix: null coordinates or weights");
  SUMPY_REQUIRE(pot || gx, "p2p_matrix: no output");
  SUMPY_REQUIRE(!gx || (gy && (dim == 2 || gz)), "p2p_matrix: incomplete gradient outputs");
  SUMPY_REQUIRE(!wx || (wy && (dim == 2 || wz)), "p2p_matrix: incomplete dipole weights");
  SUMPY_REQUIRE((tgtindices == nullptr) == (srcindices == nullptr),
                "p2p_matrix: give both index arrays or neither");
  SUMPY_REQUIRE(tgtindices || nentries == (long long)ntargets*nsources,
                "p2p_matrix: the full matrix has ntargets * nsources entries");
  if (nentries == 0) return SUMPYHIP_OK;
  MatArgs a{scale, k_re, k_im, ntargets, nsources, nentries, tx, ty, tz, sx, sy, sz,
            (const double *)w0, (const double *)wx, (const double *)wy, (const double *)wz,
            target_to_source, tgtindices, srcindices, (double *)pot, (double *)gx,
            (double *)gy, (double *)gz};
  const long long nb = (nentries + 255)/256;
  const int grid = (int)(nb < 65536 ? nb : 65536);
  hipStream_t st = (hipStream_t)stream;
  const bool dip = wx != nullptr, grad = gx != nullptr;
#define SUMPY_MAT(LAUNCH)                                                  \
  if (dip && grad) LAUNCH(true, true);                                     \
  else if (dip) LAUNCH(true, false);                                       \
  else if (grad) LAUNCH(false, true);                                      \
  else LAUNCH(false, false);
#define SUMPY_L3(D, G) p2p_matrix_real<KLAP3, D, G><<<grid, 256, 0, st>>>(a)
#define SUMPY_L2(D, G) p2p_matrix_real<KLAP2, D, G><<<grid, 256, 0, st>>>(a)
#define SUMPY_H3(D, G) p2p_matrix_helm<D, G><<<grid, 256, 0, st>>>(a)
  if (kernel_id == 3) { SUMPY_MAT(SUMPY_L3) }
  else if (kernel_id == 2) { SUMPY_MAT(SUMPY_L2) }
  else { SUMPY_MAT(SUMPY_H3) }
#undef SUMPY_L3
#undef SUMPY_L2
#undef SUMPY_H3
#undef SUMPY_MAT
  return check_launch("p2p_matrix");
}
