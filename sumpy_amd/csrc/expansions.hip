// Taylor (full and PDE-compressed) expansions of the 3D Laplace kernel:
// P2M, M2M, L2L, L2P, M2P, P2L.
//
// Reference math: sumpy/expansion/multipole.py:70-407 (multipole formation,
// evaluation, M2M), sumpy/expansion/local.py:189-493 (local formation,
// evaluation, L2L), sumpy/expansion/__init__.py:563-886 (compressed storage and
// the PDE projection P), sumpy/derivative_taker.py:195-262 (scaled Laplace
// derivatives).  The reference emits one symbolic expression per coefficient;
// here each operator is a per-box kernel that works on the full coefficient
// set in LDS and applies P (or P^T) from sparse tables:
//   P2M  c_full[mi] = sum_s w (a/rho)^mi / mi!  (a = centre - source), then
//        M = P^T c_full;
//   M2M  embed the child's stored coefficients into the full set, shift the
//        monomials one axis at a time, sum over children, M += P^T c;
//   L2L  L_full = P L, shift along each axis, select the stored ids;
//   L2P  L_full = P L folded with 1/(mi! rho^|mi|) into LDS in Horner order,
//        then a triple Horner per target (with its gradient);
//   M2P, P2L  scaled derivatives rho^|m| d^m (1/r) from the recurrence.
// Roofline: these kernels move little data (tens of MB at N = 2^20) and are
// bound by latency/VALU; they are small next to P2P and M2L.
#include "common.h"

namespace sumpyhip {
namespace {

constexpr int WAVE = 64;
constexpr int MAX_LIST_ORDER = 12;  // M2P / P2L orders (compile-time shells)

__device__ __forceinline__ int full_index(const sumpyhip_expn &e, int i, int j, int k) {
  const int p1 = e.order + 1;
  if (i < 0 || j < 0 || k < 0 || i + j + k > e.order) return -1;
  return e.cube_index[(i*p1 + j)*p1 + k];
}

__device__ __forceinline__ double inv_fact(int n) {
  double f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return 1.0/f;
}

// LDS written by some lanes of a wave and read by others of the same wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// {{{ P2M

// rscale per tree level, for launches that span several levels (a box's
// rscale is r[box_levels[box]]); unused when box_levels is null
constexpr int LIST_MAXLEV = 32;
struct LevelScales {
  double r[LIST_MAXLEV];
};

constexpr int P2M_NT = 128;   // threads per block, one source box at a time
constexpr int P2M_SCH = 32;   // sources per LDS chunk

// 1/q for q < 64 (power tables without divisions)
__constant__ double c_inv[64] = {
    0.0, 1.0, 1.0/2, 1.0/3, 1.0/4, 1.0/5, 1.0/6, 1.0/7, 1.0/8, 1.0/9, 1.0/10, 1.0/11,
    1.0/12, 1.0/13, 1.0/14, 1.0/15, 1.0/16, 1.0/17, 1.0/18, 1.0/19, 1.0/20, 1.0/21,
    1.0/22, 1.0/23, 1.0/24, 1.0/25, 1.0/26, 1.0/27, 1.0/28, 1.0/29, 1.0/30, 1.0/31,
    1.0/32, 1.0/33, 1.0/34, 1.0/35, 1.0/36, 1.0/37, 1.0/38, 1.0/39, 1.0/40, 1.0/41,
    1.0/42, 1.0/43, 1.0/44, 1.0/45, 1.0/46, 1.0/47, 1.0/48, 1.0/49, 1.0/50, 1.0/51,
    1.0/52, 1.0/53, 1.0/54, 1.0/55, 1.0/56, 1.0/57, 1.0/58, 1.0/59, 1.0/60, 1.0/61,
    1.0/62, 1.0/63};

// Number of (i, j, k-range) tasks when each pair (i, j), i + j <= p, is cut
// into pieces of at most kt consecutive k (k <= p - i - j).
__host__ __device__ inline int p2m_ntasks(int p, int kt) {
  int n = 0;
  for (int s = 0; s <= p; ++s) n += (s + 1)*((p - s + kt)/kt);
  return n;
}

// Task t: pairs in order of (i + j, i), pieces in order of k.
__device__ inline bool p2m_task(int p, int kt, int t, int &ti, int &tj, int &k0, int &k1) {
  int n = 0;
  for (int s = 0; s <= p; ++s) {
    const int pieces = (p - s + kt)/kt;
    if (t < n + (s + 1)*pieces) {
      const int q = t - n;
      ti = q/pieces;
      tj = s - ti;
      k0 = (q - ti*pieces)*kt;
      k1 = min(k0 + kt, p - s + 1);
      return true;
    }
    n += (s + 1)*pieces;
  }
  return false;
}

// C_full[i,j,k] = sum_s w0 X_i Y_j Z_k + wx X_{i-1} Y_j Z_k + wy X_i Y_{j-1} Z_k
//                 + wz X_i Y_j Z_{k-1},   X_i = (a_x/rho)^i / i!,  a = centre - source
// (sumpy/expansion/multipole.py:78-114; dipole weights pre-divided by rho),
// then M = P^T C_full.  Each thread owns an (i, j, k-range) task and keeps its
// KT sums in registers; the per-source power tables X, Y, Z live in LDS.
template <bool DIP, int KT>
__global__ __launch_bounds__(P2M_NT) void p2m_lap3d(
    sumpyhip_expn e, int nboxes_list, const int *source_boxes, const int *bss,
    const int *bsc, const double *centers, int nboxes, const double *sx,
    const double *sy, const double *sz, const double *w0, const double *wx,
    const double *wy, const double *wz, double rscale, double *out, int base,
    const int *box_levels, LevelScales ls) {
  extern __shared__ double lds[];
  const int p = e.order, P1 = p + 1;
  const int PW = P1 + KT;                 // power rows zero padded by KT
  double *pw = lds;                       // [SCH][3][PW]
  double *sw = pw + P2M_SCH*3*PW;         // [SCH][4]
  double *cf = sw + P2M_SCH*4;            // [nfull]
  const int tid = threadIdx.x;
  const int ntasks = p2m_ntasks(p, KT);
  const int nrounds = (ntasks + P2M_NT - 1)/P2M_NT;
  for (int ib = blockIdx.x; ib < nboxes_list; ib += gridDim.x) {
    const int b = source_boxes[ib];
    const double rinv = 1.0/(box_levels ? ls.r[box_levels[b]] : rscale);
    const double c0 = centers[b], c1 = centers[nboxes + b], c2 = centers[2*(size_t)nboxes + b];
    const int s0 = bss[b], ns = bsc[b];
    for (int round = 0; round < nrounds; ++round) {
      int ti = 0, tj = 0, k0 = 0, k1 = 0;
      const bool has = p2m_task(p, KT, tid + round*P2M_NT, ti, tj, k0, k1);
      double acc[KT];
      for (int kk = 0; kk < KT; ++kk) acc[kk] = 0;
      for (int sc = 0; sc < ns; sc += P2M_SCH) {
        const int nn = min(P2M_SCH, ns - sc);
        __syncthreads();
        if (tid < nn) {
          const int s = s0 + sc + tid;
          const double a[3] = {(c0 - sx[s])*rinv, (c1 - sy[s])*rinv, (c2 - sz[s])*rinv};
          double *row = pw + tid*3*PW;
          for (int d = 0; d < 3; ++d) {
            double v = 1;
            row[d*PW] = 1;
            for (int q = 1; q < P1; ++q) { v *= a[d]*c_inv[q]; row[d*PW + q] = v; }
            for (int q = P1; q < PW; ++q) row[d*PW + q] = 0.0;
          }
          sw[tid*4] = w0[s];
          if (DIP) {
            sw[tid*4 + 1] = wx[s]*rinv; sw[tid*4 + 2] = wy[s]*rinv; sw[tid*4 + 3] = wz[s]*rinv;
          }
        }
        __syncthreads();
        if (has) {
          // branch-free: the KT powers past the task's range read zero
          // padding or feed sums that are not stored
          #pragma unroll 4
          for (int s = 0; s < nn; ++s) {
            const double *X = pw + s*3*PW, *Y = X + PW, *Z = Y + PW;
            const double xi = X[ti], yj = Y[tj];
            double bco = sw[s*4]*xi*yj, cco = 0;
            if (DIP) {
              bco += sw[s*4 + 1]*(ti ? X[ti - 1] : 0.0)*yj + sw[s*4 + 2]*xi*(tj ? Y[tj - 1] : 0.0);
              cco = sw[s*4 + 3]*xi*yj;
            }
            #pragma unroll
            for (int kk = 0; kk < KT; ++kk) {
              const int k = k0 + kk;
              acc[kk] += bco*Z[k];
              if (DIP) acc[kk] += cco*(k > 0 ? Z[k - 1] : 0.0);
            }
          }
        }
      }
      if (has)
        for (int kk = 0; kk < KT; ++kk) {
          const int k = k0 + kk;
          if (k < k1) cf[e.cube_index[(ti*P1 + tj)*P1 + k]] = acc[kk];
        }
    }
    __syncthreads();
    double *o = out + (size_t)(b - base)*e.nstored;
    for (int s = tid; s < e.nstored; s += P2M_NT) {
      double v = 0;
      for (int q = e.projT_col_starts[s]; q < e.projT_col_starts[s + 1]; ++q)
        v += e.projT_vals[q]*cf[e.projT_rows[q]];
      o[s] = v;
    }
  }
}

// P2M with the order at compile time: one wave per source box (4 boxes per
// block), lane = (i, j) pair(s) of the full Taylor set, accumulating
// C[i, j, k] for all k <= ORD in registers (entries past ORD - i - j are not
// stored).  Per chunk of 32 sources the lanes tabulate X_q = (a_x/rho)^q / q!
// and Y_q in LDS (one source per lane); per source a lane gathers its X_i and
// Y_j, and runs the Z powers of the (broadcast) source itself.  The stored
// coefficients are then M = P^T C (sumpy/expansion/multipole.py:78-114).
constexpr int P2M_NW = 4;

// T = float: the source tables, the (i, j, k) moments and their sums over a
// box's sources in FP32 (box-relative scaled coordinates, |a/rho| <~ 1), the
// projection to the stored coefficients in FP64
template <int ORD, bool DIP, typename T = double>
__global__ __launch_bounds__(64*P2M_NW) void p2m_lap3d_pairs(
    sumpyhip_expn e, int nboxes_list, const int *__restrict__ source_boxes,
    const int *__restrict__ bss, const int *__restrict__ bsc,
    const double *__restrict__ centers, int nboxes, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ w0, const double *__restrict__ wx,
    const double *__restrict__ wy, const double *__restrict__ wz, double rscale, double *out,
    int base, const int *__restrict__ box_levels, LevelScales ls) {
  constexpr int P1 = ORD + 1;
  constexpr int NP = P1*(P1 + 1)/2;  // pairs i + j <= ORD
  constexpr int PPL = (NP + 63)/64;  // pairs per lane
  constexpr int NF = P1*(P1 + 1)*(P1 + 2)/6;
  constexpr int CHK = 32;            // sources per table chunk
  __shared__ T sX[P2M_NW][CHK][P1], sY[P2M_NW][CHK][P1];
  __shared__ T sZ[P2M_NW][CHK][DIP ? 4 : 2];  // a_z / rho, w0 (wx, wy / rho)
  __shared__ T sWz[P2M_NW][DIP ? CHK : 1];    // wz / rho
  __shared__ double sC[P2M_NW][NF];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // this lane's pairs, in order of i + j, then i
  int pi[PPL], pj[PPL];
  #pragma unroll
  for (int r = 0; r < PPL; ++r) {
    int q = lane + 64*r, d = 0;
    while (d <= ORD && q >= d + 1) { q -= d + 1; ++d; }
    pi[r] = d <= ORD ? q : 0;
    pj[r] = d <= ORD ? d - q : 0;
    if (d > ORD) pi[r] = -1;
  }
  for (int ib = blockIdx.x*P2M_NW + w; ib < nboxes_list; ib += gridDim.x*P2M_NW) {
    const int b = source_boxes[ib];
    const double rinv = 1.0/(box_levels ? ls.r[box_levels[b]] : rscale);
    const double c0 = centers[b], c1 = centers[nboxes + b], c2 = centers[2*(size_t)nboxes + b];
    const int s0 = bss[b], ns = bsc[b];
    T acc[PPL][P1];
    #pragma unroll
    for (int r = 0; r < PPL; ++r)
      #pragma unroll
      for (int k = 0; k < P1; ++k) acc[r][k] = T(0);
    for (int sc = 0; sc < ns; sc += CHK) {
      const int nn = min(CHK, ns - sc);
      wave_sync();
      if (lane < nn) {
        const int s = s0 + sc + lane;
        const T ax = T((c0 - sx[s])*rinv), ay = T((c1 - sy[s])*rinv);
        T xv = 1, yv = 1;
        #pragma unroll
        for (int q = 0; q < P1; ++q) {
          sX[w][lane][q] = xv;
          sY[w][lane][q] = yv;
          xv *= ax*T(1.0/(q + 1));
          yv *= ay*T(1.0/(q + 1));
        }
        sZ[w][lane][0] = T((c2 - sz[s])*rinv);
        sZ[w][lane][1] = T(w0[s]);
        if (DIP) {
          sZ[w][lane][2 % (DIP ? 4 : 2)] = T(wx[s]*rinv);
          sZ[w][lane][3 % (DIP ? 4 : 2)] = T(wy[s]*rinv);
          sWz[w][lane % (DIP ? CHK : 1)] = T(wz[s]*rinv);
        }
      }
      wave_sync();
      for (int t = 0; t < nn; ++t) {
        const T az = sZ[w][t][0], q0 = sZ[w][t][1];
        #pragma unroll
        for (int r = 0; r < PPL; ++r) {
          const int i = pi[r] < 0 ? 0 : pi[r], j = pj[r];
          const T xi = sX[w][t][i], yj = sY[w][t][j];
          T bco = q0*xi*yj, cco = T(0);
          if (DIP) {
            const T qx = sZ[w][t][2 % (DIP ? 4 : 2)], qy = sZ[w][t][3 % (DIP ? 4 : 2)];
            bco += qx*(i ? sX[w][t][i - 1] : T(0))*yj + qy*xi*(j ? sY[w][t][j - 1] : T(0));
            cco = sWz[w][t % (DIP ? CHK : 1)]*xi*yj;
          }
          T z = T(1), zprev = T(0);
          #pragma unroll
          for (int k = 0; k < P1; ++k) {
            acc[r][k] += bco*z;
            if (DIP) acc[r][k] += cco*zprev;
            zprev = z;
            z *= az*T(1.0/(k + 1));
          }
        }
      }
    }
    // C to LDS (full index order), then M = P^T C
    wave_sync();
    #pragma unroll
    for (int r = 0; r < PPL; ++r) {
      if (pi[r] < 0) continue;
      const int i = pi[r], j = pj[r];
      #pragma unroll
      for (int k = 0; k < P1; ++k)
        if (k <= ORD - i - j) sC[w][e.cube_index[(i*P1 + j)*P1 + k]] = double(acc[r][k]);
    }
    wave_sync();
    double *o = out + (size_t)(b - base)*e.nstored;
    for (int s = lane; s < e.nstored; s += 64) {
      double v = 0;
      for (int q = e.projT_col_starts[s]; q < e.projT_col_starts[s + 1]; ++q)
        v += e.projT_vals[q]*sC[w][e.projT_rows[q]];
      o[s] = v;
    }
  }
}

// smallest of the supported piece lengths that needs at most one round
inline int p2m_kt(int p) {
  const int kts[] = {2, 3, 4, 6, 12};
  for (int kt : kts)
    if (p2m_ntasks(p, kt) <= P2M_NT) return kt;
  return 12;
}

// }}}

// {{{ M2M

__global__ __launch_bounds__(WAVE) void m2m_lap3d(
    sumpyhip_expn se, sumpyhip_expn te, int ntgt, const int *target_boxes,
    const int *child_ids, const double *centers, int nboxes, double src_rscale,
    double tgt_rscale, const double *src, int sbase, double *tgt, int tbase) {
  extern __shared__ double lds[];
  const int nf = te.nfull, P1 = te.order + 1;
  double *b0 = lds, *b1 = b0 + nf, *acc = b1 + nf, *pw = acc + nf;  // pw[3][P1]
  const int lane = threadIdx.x;
  const double ratio = src_rscale/tgt_rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    for (int f = lane; f < nf; f += WAVE) acc[f] = 0;
    for (int c = 0; c < 8; ++c) {
      const int ch = child_ids[(size_t)c*nboxes + tb];
      if (ch == 0) continue;
      __syncthreads();
      for (int f = lane; f < nf; f += WAVE) b0[f] = 0;
      if (lane < 3) {
        double d = (centers[(size_t)lane*nboxes + tb] - centers[(size_t)lane*nboxes + ch])/tgt_rscale;
        double v = 1;
        pw[lane*P1] = 1;
        for (int i = 1; i < P1; ++i) { v *= d/i; pw[lane*P1 + i] = v; }
      }
      __syncthreads();
      const double *m = src + (size_t)(ch - sbase)*se.nstored;
      for (int s = lane; s < se.nstored; s += WAVE) {
        int i = se.stored_ids[s], j = se.stored_ids[se.nstored + s],
            k = se.stored_ids[2*se.nstored + s];
        int f = full_index(te, i, j, k);
        if (f >= 0) b0[f] = m[s]*pow(ratio, (double)(i + j + k));
      }
      // shift along y, z, x (sumpy/expansion/multipole.py:331-357)
      const int axes[3] = {1, 2, 0};
      double *in = b0, *outb = b1;
      for (int ax : axes) {
        __syncthreads();
        for (int f = lane; f < nf; f += WAVE) {
          int mi[3] = {te.full_ids[f], te.full_ids[nf + f], te.full_ids[2*nf + f]};
          double v = 0;
          const int n = mi[ax];
          for (int k = 0; k <= n; ++k) {
            int mk[3] = {mi[0], mi[1], mi[2]};
            mk[ax] = k;
            v += in[full_index(te, mk[0], mk[1], mk[2])]*pw[ax*P1 + n - k];
          }
          outb[f] = v;
        }
        double *t = in; in = outb; outb = t;
      }
      __syncthreads();
      for (int f = lane; f < nf; f += WAVE) acc[f] += in[f];
    }
    __syncthreads();
    double *o = tgt + (size_t)(tb - tbase)*te.nstored;
    for (int s = lane; s < te.nstored; s += WAVE) {
      double v = 0;
      for (int q = te.projT_col_starts[s]; q < te.projT_col_starts[s + 1]; ++q)
        v += te.projT_vals[q]*acc[te.projT_rows[q]];
      o[s] += v;
    }
    __syncthreads();
  }
}

// }}}

// {{{ L2L

__global__ __launch_bounds__(WAVE) void l2l_lap3d(
    sumpyhip_expn se, sumpyhip_expn te, int ntgt, const int *target_boxes,
    const int *parent_ids, const double *centers, int nboxes, double src_rscale,
    double tgt_rscale, const double *src, int sbase, double *tgt, int tbase) {
  extern __shared__ double lds[];
  const int nf = se.nfull, P1 = se.order + 1;
  double *b0 = lds, *b1 = b0 + nf, *pw = b1 + nf;
  const int lane = threadIdx.x;
  const double ratio = tgt_rscale/src_rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    const int pb = parent_ids[tb];
    __syncthreads();
    if (lane < 3) {
      double d = (centers[(size_t)lane*nboxes + tb] - centers[(size_t)lane*nboxes + pb])/src_rscale;
      double v = 1;
      pw[lane*P1] = 1;
      for (int i = 1; i < P1; ++i) { v *= d/i; pw[lane*P1 + i] = v; }
    }
    const double *L = src + (size_t)(pb - sbase)*se.nstored;
    for (int f = lane; f < nf; f += WAVE) {
      double v = 0;
      for (int q = se.proj_row_starts[f]; q < se.proj_row_starts[f + 1]; ++q)
        v += se.proj_vals[q]*L[se.proj_cols[q]];
      b0[f] = v;
    }
    // shift along x, y, z (sumpy/expansion/local.py:420-453)
    double *in = b0, *outb = b1;
    for (int ax = 0; ax < 3; ++ax) {
      __syncthreads();
      for (int f = lane; f < nf; f += WAVE) {
        int mi[3] = {se.full_ids[f], se.full_ids[nf + f], se.full_ids[2*nf + f]};
        const int deg = mi[0] + mi[1] + mi[2];
        double v = 0;
        for (int q = 0; q <= se.order - deg; ++q) {
          int mk[3] = {mi[0], mi[1], mi[2]};
          mk[ax] += q;
          v += in[full_index(se, mk[0], mk[1], mk[2])]*pw[ax*P1 + q];
        }
        outb[f] = v;
      }
      double *t = in; in = outb; outb = t;
    }
    __syncthreads();
    double *o = tgt + (size_t)(tb - tbase)*te.nstored;
    for (int s = lane; s < te.nstored; s += WAVE) {
      int i = te.stored_ids[s], j = te.stored_ids[te.nstored + s],
          k = te.stored_ids[2*te.nstored + s];
      int f = full_index(se, i, j, k);
      if (f >= 0) o[s] += in[f]*pow(ratio, (double)(i + j + k));
    }
  }
}

// }}}

// {{{ L2P

constexpr int L2P_NWMAX = 8;  // waves per block (8, or 4 where 8 exceed the LDS), each on its own target box


// Per block (once): the Horner-ordered factors h[pos] = rho^-|m| / m! of
// L_full = P L (sumpy/expansion/local.py:284-309) and P's rows in Horner
// order, in LDS.  Each wave then takes its own boxes: L (stored) to LDS,
// lh[pos] = h[pos] (P L)[m(pos)], and a triple Horner per target (one lane
// per target) with its gradient.  Every wave of a block has its own box.
// Triple Horner of the local expansion at (bx, by, bz) over the Horner-ordered
// coefficients lh (runs of n = p - i - j + 1 coefficients for i = p..0,
// j = p - i..0), with its gradient; ORD = p at compile time, so every
// coefficient read has a constant offset and no row is padded.
template <int ORD, bool GRAD, typename T>
__device__ __forceinline__ void horner_t(const T *lh, T bx, T by, T bz, T &P, T &Px, T &Py,
                                         T &Pz) {
  int pos = 0;
  #pragma unroll
  for (int i = ORD; i >= 0; --i) {
    T S = 0, Sy = 0, Sz = 0;
    #pragma unroll
    for (int j = ORD - i; j >= 0; --j) {
      const int n = ORD - i - j + 1;
      T R = 0, Rz = 0;
      #pragma unroll
      for (int q = 0; q < n; ++q) {
        if (GRAD) Rz = Rz*bz + R;
        R = R*bz + lh[pos + q];
      }
      pos += n;
      if (GRAD) { Sy = Sy*by + S; Sz = Sz*by + Rz; }
      S = S*by + R;
    }
    if (GRAD) { Px = Px*bx + P; Py = Py*bx + Sy; Pz = Pz*bx + Sz; }
    P = P*bx + S;
  }
}

// T: the type of the Horner coefficients and evaluation (double, or float
// for the mixed form; L_full = P L is formed in FP64).  The expansion is
// evaluated in b / rho (rho of the box: box_levels and ls when several levels
// share the launch), the gradient scaled back by 1 / rho.
template <bool GRAD, int KMAX, int ORD = -1, typename T = double>
__global__ __launch_bounds__(64*L2P_NWMAX) void l2p_lap3d(
    sumpyhip_expn e, int ntgt, const int *target_boxes, const int *bts,
    const int *btc, const double *centers, int nboxes, const double *tx,
    const double *ty, const double *tz, double rscale, double scale,
    const double *src, int sbase, double *pot, double *gx, double *gy, double *gz,
    const int *box_levels, LevelScales ls) {
  extern __shared__ double sh[];
  const int nf = e.nfull, nst = e.nstored, nnz = e.proj_nnz;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
  double *hs = sh;                          // [nfull] factors 1 / m!
  double *pv = hs + nf;                     // [nnz] values of P (Horner row order)
  double *Ls = pv + nnz + w*(nf + nst);     // per wave: [nstored] L of the box
  T *lh = reinterpret_cast<T *>(Ls + nst);  //           [nfull] Horner coefficients
  int *rs = reinterpret_cast<int *>(pv + nnz + nw*(nf + nst));  // [nfull + 1]
  int *pc = rs + nf + 1;                    // [nnz] columns of P
  const int p = e.order;  // p + 1 <= KMAX
  if (w == 0) {  // row starts of P in Horner order: each lane scans a contiguous range
    const int ch = (nf + 63)/64, p0 = min(nf, lane*ch), p1 = min(nf, p0 + ch);
    int loc = 0;
    for (int pos = p0; pos < p1; ++pos) {
      const int f = e.horner_order[pos];
      loc += e.proj_row_starts[f + 1] - e.proj_row_starts[f];
    }
    int incl = loc;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    int n = incl - loc;
    for (int pos = p0; pos < p1; ++pos) {
      rs[pos] = n;
      const int f = e.horner_order[pos];
      n += e.proj_row_starts[f + 1] - e.proj_row_starts[f];
    }
    if (lane == 63) rs[nf] = incl;
  }
  __syncthreads();
  for (int pos = tid; pos < nf; pos += blockDim.x) {
    const int f = e.horner_order[pos];
    const int i = e.full_ids[f], j = e.full_ids[nf + f], k = e.full_ids[2*nf + f];
    hs[pos] = inv_fact(i)*inv_fact(j)*inv_fact(k);
    int d = rs[pos];
    for (int q = e.proj_row_starts[f]; q < e.proj_row_starts[f + 1]; ++q, ++d) {
      pv[d] = e.proj_vals[q];
      pc[d] = e.proj_cols[q];
    }
  }
  __syncthreads();
  for (int ib = blockIdx.x*nw + w; ib < ntgt; ib += gridDim.x*nw) {
    const int tb = target_boxes[ib];
    const double *L = src + (size_t)(tb - sbase)*nst;
    wave_sync();
    for (int q = lane; q < nst; q += 64) Ls[q] = L[q];
    wave_sync();
    for (int pos = lane; pos < nf; pos += 64) {
      double v = 0;
      for (int q = rs[pos]; q < rs[pos + 1]; ++q) v += pv[q]*Ls[pc[q]];
      lh[pos] = T(v*hs[pos]);
    }
    wave_sync();
    const double c0 = centers[tb], c1 = centers[nboxes + tb], c2 = centers[2*(size_t)nboxes + tb];
    const double rinv = 1.0/(box_levels ? ls.r[box_levels[tb]] : rscale);
    const int t0 = bts[tb], nt = btc[tb];
    for (int tc = 0; tc < nt; tc += 64) {
      const bool valid = tc + lane < nt;
      const int t = t0 + tc + (valid ? lane : 0);
      const T bx = T((tx[t] - c0)*rinv), by = T((ty[t] - c1)*rinv), bz = T((tz[t] - c2)*rinv);
      T P = 0, Px = 0, Py = 0, Pz = 0;
      if constexpr (ORD >= 0) {
        horner_t<ORD, GRAD>(lh, bx, by, bz, P, Px, Py, Pz);
      } else {
      int pos = 0;
      for (int i = p; i >= 0; --i) {
        T S = 0, Sy = 0, Sz = 0;
        for (int j = p - i; j >= 0; --j) {
          // the (i, j) run of n coefficients: KMAX loads issued together
          const int n = p - i - j + 1;
          T c[KMAX];
          #pragma unroll
          for (int q = 0; q < KMAX; ++q) c[q] = lh[pos + (q < n ? q : n - 1)];
          T R = 0, Rz = 0;
          #pragma unroll
          for (int q = 0; q < KMAX; ++q) {
            if (q < n) {
              if (GRAD) Rz = Rz*bz + R;
              R = R*bz + c[q];
            }
          }
          pos += n;
          if (GRAD) { Sy = Sy*by + S; Sz = Sz*by + Rz; }
          S = S*by + R;
        }
        if (GRAD) { Px = Px*bx + P; Py = Py*bx + Sy; Pz = Pz*bx + Sz; }
        P = P*bx + S;
      }
      }
      if (valid) {
        if (pot) pot[t] = scale*double(P);
        if (GRAD) {
          const double gs = scale*rinv;
          gx[t] = gs*double(Px); gy[t] = gs*double(Py); gz[t] = gs*double(Pz);
        }
      }
    }
  }
}

// }}}

// {{{ scaled Laplace derivatives for M2P/P2L, one shell |m| = n at a time

// The values D(m) = rho^|m| d^m (1/r), m_x <= 1, |m| <= Q, are produced by
// the recurrence of sumpy/derivative_taker.py:227-259 shell by shell: shell n
// (the 2n + 1 values with |m| = n) needs only shells n - 1 and n - 2, so with
// the order a template parameter the three live shells stay in registers and
// every index is a compile-time constant.  Shell n holds m at
// m_x ? n + 1 + m_y : m_y, and the concatenation of shells 0..n puts m at
// n^2 + that index (the PDE-compressed stored set is exactly m_x <= 1).

__host__ __device__ constexpr int sidx(int n, int mx, int my) { return mx ? n + 1 + my : my; }

__device__ __forceinline__ int shell_pos(int i, int j, int k) {
  const int n = i + j + k;
  return n*n + sidx(n, i, j);
}

// T: the arithmetic type of the recurrence (double, or float for the mixed
// form's list-3/4 terms; the offset x is formed in FP64 by the caller)
template <typename T = double>
struct LapGeo {
  T xs[3];     // x / rho
  T ri2, rinv; // 1 / |x / rho|^2, 1 / |x|
  __device__ LapGeo(double x, double y, double z, double rscale) {
    const double inv = 1.0/rscale;
    xs[0] = T(x*inv); xs[1] = T(y*inv); xs[2] = T(z*inv);
    const T r2s = xs[0]*xs[0] + xs[1]*xs[1] + xs[2]*xs[2];
    const T ri = rsqrt(r2s);
    ri2 = ri*ri;
    rinv = ri*T(inv);
  }
};

// shell N from shells N - 1 (p1) and N - 2 (p2)
template <int N, typename T>
__device__ __forceinline__ void lap3d_shell(const LapGeo<T> &g, const T *p1, const T *p2,
                                            T *cur) {
  #pragma unroll
  for (int mx = 0; mx <= 1; ++mx) {
    #pragma unroll
    for (int my = 0; my <= N - mx; ++my) {
      const int mz = N - mx - my;
      T val;
      if (my <= 1 && mz <= 1) {
        // m in {0,1}^3: closed forms
        if (N == 0) {
          val = g.rinv;
        } else if (N == 1) {
          val = -g.xs[mx ? 0 : (my ? 1 : 2)]*g.ri2*g.rinv;
        } else if (N == 2) {
          const int a = mx ? 0 : 1, b = mz ? 2 : 1;
          val = T(3)*g.xs[a]*g.xs[b]*(g.ri2*g.ri2)*g.rinv;
        } else {
          val = T(-15)*g.xs[0]*g.xs[1]*g.xs[2]*(g.ri2*g.ri2*g.ri2)*g.rinv;
        }
      } else {
        // d: the first axis with m_d >= 2 (m_x <= 1, so y or z)
        const int d = my >= 2 ? 1 : 2;
        T acc = 0;
        if (mx == 1) acc += T(2)*g.xs[0]*p1[sidx(N - 1, 0, my)];
        if (my >= 1) {
          const T c1 = d == 1 ? 2*my - 1 : 2*my;
          const T c2 = d == 1 ? (my - 1)*(my - 1) : my*(my - 1);
          acc += c1*g.xs[1]*p1[sidx(N - 1, mx, my - 1)];
          if (my >= 2) acc += c2*p2[sidx(N - 2, mx, my - 2)];
        }
        if (mz >= 1) {
          const T c1 = d == 2 ? 2*mz - 1 : 2*mz;
          const T c2 = d == 2 ? (mz - 1)*(mz - 1) : mz*(mz - 1);
          acc += c1*g.xs[2]*p1[sidx(N - 1, mx, my)];
          if (mz >= 2) acc += c2*p2[sidx(N - 2, mx, my)];
        }
        val = -acc*g.ri2;
      }
      cur[sidx(N, mx, my)] = val;
    }
  }
}

// Runs shells N..Q, handing each to f.template shell<N>(cur, prev) (prev =
// shell N - 1).
template <int N, int Q, bool DONE = (N > Q)>
struct LapShells {
  template <typename T, class F>
  static __device__ __forceinline__ void run(const LapGeo<T> &g, const T *p2, const T *p1,
                                             F &f) {
    T cur[2*N + 1];
    lap3d_shell<N>(g, p1, p2, cur);
    f.template shell<N>(cur, p1);
    LapShells<N + 1, Q>::run(g, p1, cur, f);
  }
};

template <int N, int Q>
struct LapShells<N, Q, true> {
  template <typename T, class F>
  static __device__ __forceinline__ void run(const LapGeo<T> &, const T *, const T *, F &) {}
};

template <int Q, typename T, class F>
__device__ __forceinline__ void lap3d_shells(const LapGeo<T> &g, F &f) {
  const T none[1] = {T(0)};
  LapShells<0, Q>::run(g, none, none, f);
}

// D(m + e_x) for m = (mx, my, .) of shell N - 1, from shell N (m_x = 2 by
// Laplace's equation: D(2, y, z) = -D(0, y + 2, z) - D(0, y, z + 2))
template <int N, typename T>
__device__ __forceinline__ T shift_x(const T *cur, int mx, int my) {
  return mx == 0 ? cur[sidx(N, 1, my)] : -cur[sidx(N, 0, my + 2)] - cur[sidx(N, 0, my)];
}

// }}}

// {{{ M2P (list 3), accumulate

constexpr int LW = 4;  // waves per block of the list 3 / 4 kernels



// M2P consumer: P += sum_m M(m) D(m) (shells <= ORD); with GRAD also
// G_ax += sum_m M(m) D(m + e_ax) from the next shell.
template <int ORD, bool GRAD, typename T = double>
struct M2PSum {
  const T *M;       // this lane's source box, shell order (LDS)
  T gs = 1;         // 1 / rscale of that box (the gradient's extra derivative)
  T P = 0, G[3] = {0, 0, 0};                // this box's sums (type T)
  double Pt = 0, Gt[3] = {0, 0, 0};         // over the boxes (FP64)
  __device__ __forceinline__ void flush() {
    Pt += double(P);
    P = 0;
    if (GRAD)
      for (int c = 0; c < 3; ++c) { Gt[c] += double(G[c]); G[c] = 0; }
  }
  template <int N>
  __device__ __forceinline__ void shell(const T *cur, const T *prev) {
    if (N <= ORD) {
      #pragma unroll
      for (int i = 0; i < 2*N + 1; ++i) P += M[N*N + i]*cur[i];
    }
    if (GRAD && N >= 1) {
      constexpr int n = N - 1;
      #pragma unroll
      for (int mx = 0; mx <= 1; ++mx) {
        #pragma unroll
        for (int my = 0; my <= n - mx; ++my) {
          const T m = gs*M[n*n + sidx(n, mx, my)];
          G[0] += m*shift_x<N>(cur, mx, my);
          G[1] += m*cur[sidx(N, mx, my + 1)];
          G[2] += m*cur[sidx(N, mx, my)];
        }
      }
    }
  }
};

// One block of LW waves per target box.  Lanes are targets; when the box has
// at most 32 (16) targets a wave's lanes form 2 (4) groups of 32 (16), each on
// its own source box.  Every wave stages the multipoles of its current source
// boxes in LDS in shell order; the partial sums of waves and groups are added
// in LDS at the end.
template <int ORD, bool GRAD, typename T = double>
__global__ __launch_bounds__(64*LW) void m2p_lap3d(
    sumpyhip_expn e, int ntgt, const int *__restrict__ target_boxes,
    const int *__restrict__ sstarts, const int *__restrict__ slists,
    const int *__restrict__ bts, const int *__restrict__ btc,
    const double *__restrict__ centers, int nboxes, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz, double rscale, double scale,
    const double *__restrict__ src, int sbase, double *pot, double *gx, double *gy,
    double *gz, const int *__restrict__ box_levels, LevelScales ls) {
  constexpr int NS = (ORD + 1)*(ORD + 1);
  constexpr int NC = GRAD ? 4 : 1;
  __shared__ int sPos[NS];
  __shared__ T sM[LW][4][NS];
  __shared__ double red[NC][LW][64];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int s = tid; s < NS; s += 64*LW)
    sPos[s] = shell_pos(e.stored_ids[s], e.stored_ids[NS + s], e.stored_ids[2*NS + s]);
  __syncthreads();
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int is0 = sstarts[ib], is1 = sstarts[ib + 1];
    if (is0 == is1) continue;
    const int tb = target_boxes[ib];
    const int t0 = bts[tb], nt = btc[tb];
    const int G = nt <= 16 ? 16 : nt <= 32 ? 32 : 64, ngrp = 64/G;
    const int grp = lane/G, tl = lane - grp*G;
    for (int tc = 0; tc < nt; tc += G) {
      const int t = t0 + tc + (tc + tl < nt ? tl : 0);
      const double x = tx[t], y = ty[t], z = tz[t];
      M2PSum<ORD, GRAD, T> f;
      f.M = sM[wave][grp];
      for (int base = is0 + wave*ngrp; base < is1; base += LW*ngrp) {
        wave_sync();
        for (int v = lane; v < ngrp*NS; v += 64) {
          const int g2 = v/NS, s = v - g2*NS;
          if (base + g2 < is1)
            sM[wave][g2][sPos[s]] = T(src[(size_t)(slists[base + g2] - sbase)*NS + s]);
        }
        wave_sync();
        if (base + grp < is1) {
          const int sb = slists[base + grp];
          const double rs = box_levels ? ls.r[box_levels[sb]] : rscale;
          f.gs = T(1.0/rs);
          const LapGeo<T> g(x - centers[sb], y - centers[nboxes + sb],
                            z - centers[2*(size_t)nboxes + sb], rs);
          lap3d_shells<ORD + (GRAD ? 1 : 0)>(g, f);
          f.flush();
        }
      }
      red[0][wave][lane] = f.Pt;
      if (GRAD)
        for (int c = 0; c < 3; ++c) red[1 + c][wave][lane] = f.Gt[c];
      __syncthreads();
      if (tid < G && tc + tid < nt) {
        double v[NC];
        for (int c = 0; c < NC; ++c) {
          v[c] = 0;
          for (int w = 0; w < LW; ++w)
            for (int g2 = 0; g2 < ngrp; ++g2) v[c] += red[c][w][g2*G + tid];
        }
        const int tt = t0 + tc + tid;
        if (pot) pot[tt] += scale*v[0];
        if (GRAD) {
          gx[tt] += scale*v[1]; gy[tt] += scale*v[2]; gz[tt] += scale*v[3];
        }
      }
      __syncthreads();
    }
  }
}

// }}}

// {{{ P2L (list 4), overwrite

// P2L consumer: every lane holds one source; the values q0 D(m) (+ sum_ax
// qd_ax D(m + e_ax) for dipoles) are produced in shell order, written to the
// wave's LDS ring (one row of 64 lanes per coefficient, two chunks of 16
// rows) and, every 16 coefficients, summed in quarters: lane l adds lanes
// 16 (l % 4) .. 16 (l % 4) + 15 of coefficient 16 k + l / 4 to its register
// acc[k].  The four quarters are added once per target box (finish()).
template <int ORD, bool DIP, typename T = double>
struct P2LSum {
  static constexpr int NS = (ORD + 1)*(ORD + 1);
  static constexpr int NCH = (NS + 15)/16;
  static constexpr int BS = 64 + 1;  // LDS row stride (values)
  static constexpr int RING = 32;    // rows: two chunks
  double acc[NCH];                   // FP64 over the chunks of sources
  T q0, qd[3];
  T *buf;  // [RING][BS]
  int lane;
  __device__ __forceinline__ void put(int p, T v) {
    buf[(p & (RING - 1))*BS + lane] = v;
    if ((p & 15) == 15 || p == NS - 1) {
      // the other chunk's rows were read before this chunk's first write
      // (previous flush), so one wave barrier per flush suffices
      wave_sync();
      const T *r = buf + ((p & 16) + (lane >> 2))*BS + (lane & 3)*16;
      T t = 0;
      #pragma unroll
      for (int q = 0; q < 16; ++q) t += r[q];
      acc[p >> 4] += double(t);
    }
  }
  __device__ __forceinline__ double finish(int k) const {
    double t = acc[k];
    t += __shfl_xor(t, 1, 4);
    t += __shfl_xor(t, 2, 4);
    return t;
  }
  template <int N>
  __device__ __forceinline__ void shell(const T *cur, const T *prev) {
    if (!DIP && N <= ORD) {
      #pragma unroll
      for (int i = 0; i < 2*N + 1; ++i) put(N*N + i, q0*cur[i]);
    }
    if (DIP && N >= 1) {
      constexpr int n = N - 1;
      #pragma unroll
      for (int mx = 0; mx <= 1; ++mx) {
        #pragma unroll
        for (int my = 0; my <= n - mx; ++my) {
          const int i = sidx(n, mx, my);
          put(n*n + i, q0*prev[i] + qd[0]*shift_x<N>(cur, mx, my) +
                           qd[1]*cur[sidx(N, mx, my + 1)] + qd[2]*cur[sidx(N, mx, my)]);
        }
      }
    }
  }
};

// One wave per target box (P2L_LW boxes per block).  The sources of the box's
// list-4 boxes are concatenated (box offsets and first sources from a wave
// scan, 64 boxes at a time, in LDS) and taken 64 at a time, one per lane;
// lanes past the end carry zero weights.  The lane sums are reduced through
// LDS as the coefficients are produced (P2LSum), so a lane needs only the
// live derivative shells and (ORD + 1)^2 / 16 accumulators.
constexpr int P2L_LW = 2;  // waves (target boxes) per block of the P2L kernel

template <int ORD, bool DIP, typename T = double>
__global__ __launch_bounds__(64*P2L_LW) void p2l_lap3d(
    sumpyhip_expn e, int ntgt, const int *__restrict__ target_boxes,
    const int *__restrict__ sstarts, const int *__restrict__ slists,
    const int *__restrict__ bss, const int *__restrict__ bsc,
    const double *__restrict__ centers, int nboxes, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ w0, const double *__restrict__ wx,
    const double *__restrict__ wy, const double *__restrict__ wz, double rscale_level,
    double *out, int base, const int *__restrict__ box_levels, LevelScales ls) {
  using Sum = P2LSum<ORD, DIP, T>;
  constexpr int NS = Sum::NS;
  __shared__ int sStored[NS];  // shell position -> stored index
  __shared__ int sOff[P2L_LW][64], sFirst[P2L_LW][64];
  __shared__ T sBuf[P2L_LW][Sum::RING*Sum::BS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int s = tid; s < NS; s += 64*P2L_LW)
    sStored[shell_pos(e.stored_ids[s], e.stored_ids[NS + s], e.stored_ids[2*NS + s])] = s;
  __syncthreads();
  int *off = sOff[wave], *first = sFirst[wave];
  for (int ib = blockIdx.x*P2L_LW + wave; ib < ntgt; ib += gridDim.x*P2L_LW) {
    const int tb = target_boxes[ib];
    const int is0 = sstarts[ib], is1 = sstarts[ib + 1];
    const double rscale = box_levels ? ls.r[box_levels[tb]] : rscale_level;
    double *o = out + (size_t)(tb - base)*NS;
    if (is0 == is1) {
      // empty list 4: the expansion is overwritten with zeros
      for (int m = lane; m < NS; m += 64) o[m] = 0.0;
      continue;
    }
    const double c0 = centers[tb], c1 = centers[nboxes + tb], c2 = centers[2*(size_t)nboxes + tb];
    Sum f;
    f.buf = sBuf[wave];
    f.lane = lane;
    #pragma unroll
    for (int k = 0; k < Sum::NCH; ++k) f.acc[k] = 0.0;
    for (int cb = is0; cb < is1; cb += 64) {
      const int nb = min(64, is1 - cb);
      // exclusive scan of the source counts of boxes cb .. cb + nb - 1
      const int sb = lane < nb ? slists[cb + lane] : 0;
      const int c = lane < nb ? bsc[sb] : 0;
      int v = c;
      for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
      }
      const int total = __shfl(v, 63, 64);
      wave_sync();
      if (lane < nb) { off[lane] = v - c; first[lane] = bss[sb]; }
      wave_sync();
      // source of concatenated position j (clamped to the last one)
      auto source_at = [&](int j) {
        const int jj = min(j, total - 1);
        int lo = 0, hi = nb - 1;  // largest b with off[b] <= jj
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (off[mid] <= jj) lo = mid; else hi = mid - 1;
        }
        return first[lo] + (jj - off[lo]);
      };
      // the next batch's source data are loaded while this batch's shells run
      double nx, ny, nz, nq, ndx = 0, ndy = 0, ndz = 0;
      auto load = [&](int s) {
        nx = sx[s]; ny = sy[s]; nz = sz[s]; nq = w0[s];
        if (DIP) { ndx = wx[s]; ndy = wy[s]; ndz = wz[s]; }
      };
      if (total > 0) load(source_at(lane));
      for (int j0 = 0; j0 < total; j0 += 64) {
        const bool live = j0 + lane < total;
        const double x = nx, y = ny, z = nz;
        f.q0 = live ? T(nq) : T(0);
        if (DIP) {
          f.qd[0] = live ? T(ndx/rscale) : T(0);
          f.qd[1] = live ? T(ndy/rscale) : T(0);
          f.qd[2] = live ? T(ndz/rscale) : T(0);
        }
        load(source_at(j0 + 64 + lane));
        const LapGeo<T> g(c0 - x, c1 - y, c2 - z, rscale);
        lap3d_shells<ORD + (DIP ? 1 : 0)>(g, f);
      }
    }
    #pragma unroll
    for (int k = 0; k < Sum::NCH; ++k) {
      const int m = 16*k + (lane >> 2);
      const double v = f.finish(k);
      if ((lane & 3) == 0 && m < NS) o[sStored[m]] = v;
    }
  }
}

// }}}

int grid_for(int n) {
  const int cap = 32*num_cus();
  return n < cap ? n : cap;
}

int check_expn(const sumpyhip_expn *e, const char *what) {
  SUMPY_REQUIRE(e != nullptr, "%s: null expansion tables", what);
  SUMPY_REQUIRE(e->order >= 0 && e->order <= 24, "%s: order %d out of range", what, e->order);
  SUMPY_REQUIRE(e->nfull > 0 && e->nstored > 0 && e->full_ids && e->stored_ids &&
                    e->cube_index && e->proj_row_starts && e->projT_col_starts &&
                    e->horner_order,
                "%s: incomplete expansion tables", what);
  return SUMPYHIP_OK;
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

static int p2m_impl(int kernel_id, const sumpyhip_expn *expn, int nsrc_boxes,
                    const int *source_boxes, const int *box_source_starts,
                    const int *box_source_counts_nonchild, const double *box_centers,
                    int nboxes, const double *sx, const double *sy, const double *sz,
                    const double *w0, const double *wx, const double *wy, const double *wz,
                    double rscale, double *tgt_expansions, int tgt_base_ibox,
                    const int *box_levels, const LevelScales &ls, void *stream,
                    int precision = 0) {
  // Helmholtz (13): one real plane of complex coefficients, with projection
  // tables built for (k rscale)^2 -- the kernel itself is table-driven
  SUMPY_REQUIRE(kernel_id == 3 || kernel_id == 13,
                "p2m: only the Laplace and Helmholtz 3D kernels are implemented");
  int rc = check_expn(expn, "p2m");
  if (rc) return rc;
  if (nsrc_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(!wx || (wy && wz), "p2m: incomplete dipole weights");
  const int kt = p2m_kt(expn->order);
  const size_t lds = sizeof(double)*(3*P2M_SCH*(expn->order + 1 + kt) + 4*P2M_SCH + expn->nfull);
  SUMPY_REQUIRE(lds <= 64*1024, "p2m: order too large");
  hipStream_t st = (hipStream_t)stream;
  const int grid = nsrc_boxes;  // one box per block: many boxes in flight per CU
#define SUMPY_P2M(KT)                                                                     \
  do {                                                                                   \
    if (wx)                                                                              \
      p2m_lap3d<true, KT><<<grid, P2M_NT, lds, st>>>(                                    \
          *expn, nsrc_boxes, source_boxes, box_source_starts, box_source_counts_nonchild, \
          box_centers, nboxes, sx, sy, sz, w0, wx, wy, wz, rscale, tgt_expansions,       \
          tgt_base_ibox, box_levels, ls);                                                \
    else                                                                                 \
      p2m_lap3d<false, KT><<<grid, P2M_NT, lds, st>>>(                                   \
          *expn, nsrc_boxes, source_boxes, box_source_starts, box_source_counts_nonchild, \
          box_centers, nboxes, sx, sy, sz, w0, wx, wy, wz, rscale, tgt_expansions,       \
          tgt_base_ibox, box_levels, ls);                                                \
  } while (0)
  if (expn->nfull == (expn->order + 1)*(expn->order + 2)*(expn->order + 3)/6) {
    // pair kernel for the usual orders
    const int g2 = (nsrc_boxes + P2M_NW - 1)/P2M_NW;
#define SUMPY_P2M_PAIRS_T(O, T)                                                              \
    if (wx)                                                                                  \
      p2m_lap3d_pairs<O, true, T><<<g2, 64*P2M_NW, 0, st>>>(                                 \
          *expn, nsrc_boxes, source_boxes, box_source_starts, box_source_counts_nonchild,     \
          box_centers, nboxes, sx, sy, sz, w0, wx, wy, wz, rscale, tgt_expansions,           \
          tgt_base_ibox, box_levels, ls);                                                    \
    else                                                                                     \
      p2m_lap3d_pairs<O, false, T><<<g2, 64*P2M_NW, 0, st>>>(                                \
          *expn, nsrc_boxes, source_boxes, box_source_starts, box_source_counts_nonchild,     \
          box_centers, nboxes, sx, sy, sz, w0, wx, wy, wz, rscale, tgt_expansions,           \
          tgt_base_ibox, box_levels, ls);
#define SUMPY_P2M_PAIRS(O)                                                                   \
  case O:                                                                                    \
    if (precision == 1) {                                                                    \
      SUMPY_P2M_PAIRS_T(O, float)                                                            \
    } else {                                                                                 \
      SUMPY_P2M_PAIRS_T(O, double)                                                           \
    }                                                                                        \
    return check_launch("p2m");
    switch (expn->order) {
      SUMPY_P2M_PAIRS(1) SUMPY_P2M_PAIRS(2) SUMPY_P2M_PAIRS(3) SUMPY_P2M_PAIRS(4)
      SUMPY_P2M_PAIRS(5) SUMPY_P2M_PAIRS(6) SUMPY_P2M_PAIRS(7) SUMPY_P2M_PAIRS(8)
      SUMPY_P2M_PAIRS(9) SUMPY_P2M_PAIRS(10) SUMPY_P2M_PAIRS(11) SUMPY_P2M_PAIRS(12)
      default: break;
    }
#undef SUMPY_P2M_PAIRS
#undef SUMPY_P2M_PAIRS_T
  }
  switch (kt) {
    case 2: SUMPY_P2M(2); break;
    case 3: SUMPY_P2M(3); break;
    case 4: SUMPY_P2M(4); break;
    case 6: SUMPY_P2M(6); break;
    default: SUMPY_P2M(12); break;
  }
#undef SUMPY_P2M
  return check_launch("p2m");
}

extern "C" int sumpyhip_m2m(int kernel_id, const sumpyhip_expn *src_expn,
                            const sumpyhip_expn *tgt_expn, int ntgt_boxes,
                            const int *target_boxes, const int *box_child_ids,
                            const double *box_centers, int nboxes, double src_rscale,
                            double tgt_rscale, const double *src_expansions,
                            int src_base_ibox, double *tgt_expansions, int tgt_base_ibox,
                            void *stream) {
  // Helmholtz (13): one real plane of complex coefficients, with projection
  // tables built for (k rscale)^2 -- the kernel itself is table-driven
  SUMPY_REQUIRE(kernel_id == 3 || kernel_id == 13,
                "m2m: only the Laplace and Helmholtz 3D kernels are implemented");
  int rc = check_expn(src_expn, "m2m");
  if (!rc) rc = check_expn(tgt_expn, "m2m");
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  const size_t lds = sizeof(double)*(3*tgt_expn->nfull + 3*(tgt_expn->order + 1));
  SUMPY_REQUIRE(lds <= 64*1024, "m2m: order too large");
  m2m_lap3d<<<grid_for(ntgt_boxes), 64, lds, (hipStream_t)stream>>>(
      *src_expn, *tgt_expn, ntgt_boxes, target_boxes, box_child_ids, box_centers, nboxes,
      src_rscale, tgt_rscale, src_expansions, src_base_ibox, tgt_expansions, tgt_base_ibox);
  return check_launch("m2m");
}

extern "C" int sumpyhip_l2l(int kernel_id, const sumpyhip_expn *src_expn,
                            const sumpyhip_expn *tgt_expn, int ntgt_boxes,
                            const int *target_boxes, const int *box_parent_ids,
                            const double *box_centers, int nboxes, double src_rscale,
                            double tgt_rscale, const double *src_expansions,
                            int src_base_ibox, double *tgt_expansions, int tgt_base_ibox,
                            void *stream) {
  // Helmholtz (13): one real plane of complex coefficients, with projection
  // tables built for (k rscale)^2 -- the kernel itself is table-driven
  SUMPY_REQUIRE(kernel_id == 3 || kernel_id == 13,
                "l2l: only the Laplace and Helmholtz 3D kernels are implemented");
  int rc = check_expn(src_expn, "l2l");
  if (!rc) rc = check_expn(tgt_expn, "l2l");
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  const size_t lds = sizeof(double)*(2*src_expn->nfull + 3*(src_expn->order + 1));
  SUMPY_REQUIRE(lds <= 64*1024, "l2l: order too large");
  l2l_lap3d<<<grid_for(ntgt_boxes), 64, lds, (hipStream_t)stream>>>(
      *src_expn, *tgt_expn, ntgt_boxes, target_boxes, box_parent_ids, box_centers, nboxes,
      src_rscale, tgt_rscale, src_expansions, src_base_ibox, tgt_expansions, tgt_base_ibox);
  return check_launch("l2l");
}

static int l2p_impl(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                    const int *target_boxes, const int *box_target_starts,
                    const int *box_target_counts_nonchild, const double *box_centers,
                    int nboxes, const double *tx, const double *ty, const double *tz,
                    double rscale, double scale, const double *src_expansions,
                    int src_base_ibox, double *pot, double *gx, double *gy, double *gz,
                    const int *box_levels, const LevelScales &ls, int precision,
                    void *stream) {
  SUMPY_REQUIRE(precision == 0 || precision == 1, "l2p: precision must be 0 or 1");
  // Helmholtz (13): one real plane of complex coefficients, with projection
  // tables built for (k rscale)^2 -- the kernel itself is table-driven
  SUMPY_REQUIRE(kernel_id == 3 || kernel_id == 13,
                "l2p: only the Laplace and Helmholtz 3D kernels are implemented");
  int rc = check_expn(expn, "l2p");
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(pot || gx, "l2p: no output");
  SUMPY_REQUIRE(!gx || (gy && gz), "l2p: incomplete gradient outputs");
  SUMPY_REQUIRE(expn->proj_nnz > 0, "l2p: empty projection");
  // P (values, columns, row starts) and the factors are staged once per
  // block and shared by its waves: 8 waves per block where that fits in 64 KB
  auto lds_for = [&](int nw) {
    return sizeof(double)*(expn->nfull + expn->proj_nnz +
                           nw*(size_t)(expn->nfull + expn->nstored)) +
           sizeof(int)*(expn->nfull + 1 + (size_t)expn->proj_nnz);
  };
  const int nw = lds_for(L2P_NWMAX) <= 64*1024 ? L2P_NWMAX : 4;
  const size_t lds = lds_for(nw);
  SUMPY_REQUIRE(lds <= 64*1024, "l2p: order too large");
  const int grid = (ntgt_boxes + nw - 1)/nw < 8*num_cus() ? (ntgt_boxes + nw - 1)/nw
                                                          : 8*num_cus();
  hipStream_t st = (hipStream_t)stream;
#define SUMPY_L2P_LAUNCH_T(G, K, O, T)                                                     \
  l2p_lap3d<G, K, O, T><<<grid, 64*nw, lds, st>>>(                                       \
      *expn, ntgt_boxes, target_boxes, box_target_starts, box_target_counts_nonchild,      \
      box_centers, nboxes, tx, ty, tz, rscale, scale, src_expansions, src_base_ibox, pot,  \
      gx, gy, gz, box_levels, ls)
#define SUMPY_L2P_ORD(O)                                                                    \
  case O:                                                                                   \
    if (precision) SUMPY_L2P_LAUNCH_T(false, 8, O, float);                                  \
    else SUMPY_L2P_LAUNCH_T(false, 8, O, double);                                           \
    return check_launch("l2p");
  // potentials: compile-time Horner for the usual orders (with gradients the
  // unrolled form would hold every coefficient in registers); otherwise the
  // generic loop
  if (!gx) switch (expn->order) {
    SUMPY_L2P_ORD(1) SUMPY_L2P_ORD(2) SUMPY_L2P_ORD(3) SUMPY_L2P_ORD(4) SUMPY_L2P_ORD(5)
    SUMPY_L2P_ORD(6) SUMPY_L2P_ORD(7) SUMPY_L2P_ORD(8) SUMPY_L2P_ORD(9) SUMPY_L2P_ORD(10)
    SUMPY_L2P_ORD(11) SUMPY_L2P_ORD(12)
    default: break;
  }
#undef SUMPY_L2P_ORD
  // the run-time-order loop: FP64 only
  SUMPY_REQUIRE(precision == 0 || (!gx && expn->order >= 1 && expn->order <= 12),
                "l2p: FP32 evaluation is built for potentials of orders 1..12");
  if (expn->order < 8) {
    if (gx) SUMPY_L2P_LAUNCH_T(true, 8, -1, double);
    else SUMPY_L2P_LAUNCH_T(false, 8, -1, double);
  } else if (expn->order < 12) {
    if (gx) SUMPY_L2P_LAUNCH_T(true, 12, -1, double);
    else SUMPY_L2P_LAUNCH_T(false, 12, -1, double);
  } else {
    if (gx) SUMPY_L2P_LAUNCH_T(true, 25, -1, double);
    else SUMPY_L2P_LAUNCH_T(false, 25, -1, double);
  }
#undef SUMPY_L2P_LAUNCH_T
  return check_launch("l2p");
}

static int m2p_impl(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                    const int *target_boxes, const int *source_box_starts,
                    const int *source_box_lists, const int *box_target_starts,
                    const int *box_target_counts_nonchild, const double *box_centers,
                    int nboxes, const double *tx, const double *ty, const double *tz,
                    double rscale, double scale, const double *src_expansions,
                    int src_base_ibox, double *pot, double *gx, double *gy, double *gz,
                    const int *box_levels, const LevelScales &ls, int precision,
                    void *stream) {
  SUMPY_REQUIRE(kernel_id == 3, "m2p: only the Laplace 3D kernel is implemented");
  SUMPY_REQUIRE(precision == 0 || precision == 1, "m2p: precision must be 0 or 1");
  int rc = check_expn(expn, "m2p");
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(expn->order <= MAX_LIST_ORDER, "m2p: order %d > %d", expn->order,
                MAX_LIST_ORDER);
  SUMPY_REQUIRE(expn->nstored == (expn->order + 1)*(expn->order + 1),
                "m2p: only PDE-compressed expansions are implemented");
  SUMPY_REQUIRE(!gx || (gy && gz), "m2p: incomplete gradient outputs");
  const int grid = ntgt_boxes < (1 << 30) ? ntgt_boxes : (1 << 30);
#define SUMPY_M2P_T(ORD, GRAD, T)                                                            \
  m2p_lap3d<ORD, GRAD, T><<<grid, 64*LW, 0, (hipStream_t)stream>>>(                          \
      *expn, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,                  \
      box_target_starts, box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz,        \
      rscale, scale, src_expansions, src_base_ibox, pot, gx, gy, gz, box_levels, ls)
#define SUMPY_M2P(ORD)                                                                       \
  case ORD:                                                                                  \
    if (gx) {                                                                                \
      if (precision) SUMPY_M2P_T(ORD, true, float); else SUMPY_M2P_T(ORD, true, double);    \
    } else {                                                                                 \
      if (precision) SUMPY_M2P_T(ORD, false, float); else SUMPY_M2P_T(ORD, false, double);  \
    }                                                                                        \
    break;
  switch (expn->order) {
    SUMPY_M2P(0) SUMPY_M2P(1) SUMPY_M2P(2) SUMPY_M2P(3) SUMPY_M2P(4) SUMPY_M2P(5)
    SUMPY_M2P(6) SUMPY_M2P(7) SUMPY_M2P(8) SUMPY_M2P(9) SUMPY_M2P(10) SUMPY_M2P(11)
    SUMPY_M2P(12)
  }
#undef SUMPY_M2P
#undef SUMPY_M2P_T
  return check_launch("m2p");
}

static int p2l_impl(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                    const int *target_boxes, const int *source_box_starts,
                    const int *source_box_lists, const int *box_source_starts,
                    const int *box_source_counts_nonchild, const double *box_centers,
                    int nboxes, const double *sx, const double *sy, const double *sz,
                    const double *w0, const double *wx, const double *wy, const double *wz,
                    double rscale, double *tgt_expansions, int tgt_base_ibox,
                    const int *box_levels, const LevelScales &ls, int precision,
                    void *stream) {
  SUMPY_REQUIRE(kernel_id == 3, "p2l: only the Laplace 3D kernel is implemented");
  SUMPY_REQUIRE(precision == 0 || precision == 1, "p2l: precision must be 0 or 1");
  int rc = check_expn(expn, "p2l");
  if (rc) return rc;
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(expn->order <= MAX_LIST_ORDER, "p2l: order %d > %d", expn->order,
                MAX_LIST_ORDER);
  SUMPY_REQUIRE(expn->nstored == (expn->order + 1)*(expn->order + 1),
                "p2l: only PDE-compressed expansions are implemented");
  SUMPY_REQUIRE(!wx || (wy && wz), "p2l: incomplete dipole weights");
  const int grid = (ntgt_boxes + P2L_LW - 1)/P2L_LW;
#define SUMPY_P2L_T(ORD, DIP, T)                                                             \
  p2l_lap3d<ORD, DIP, T><<<grid, 64*P2L_LW, 0, (hipStream_t)stream>>>(                       \
      *expn, ntgt_boxes, target_boxes, source_box_starts, source_box_lists,                  \
      box_source_starts, box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz, w0,    \
      wx, wy, wz, rscale, tgt_expansions, tgt_base_ibox, box_levels, ls)
#define SUMPY_P2L(ORD)                                                                       \
  case ORD:                                                                                  \
    if (wx) {                                                                                \
      if (precision) SUMPY_P2L_T(ORD, true, float); else SUMPY_P2L_T(ORD, true, double);    \
    } else {                                                                                 \
      if (precision) SUMPY_P2L_T(ORD, false, float); else SUMPY_P2L_T(ORD, false, double);  \
    }                                                                                        \
    break;
  switch (expn->order) {
    SUMPY_P2L(0) SUMPY_P2L(1) SUMPY_P2L(2) SUMPY_P2L(3) SUMPY_P2L(4) SUMPY_P2L(5)
    SUMPY_P2L(6) SUMPY_P2L(7) SUMPY_P2L(8) SUMPY_P2L(9) SUMPY_P2L(10) SUMPY_P2L(11)
    SUMPY_P2L(12)
  }
#undef SUMPY_P2L
#undef SUMPY_P2L_T
  return check_launch("p2l");
}

static int level_scales(int nlevels, const double *level_rscale, LevelScales &ls,
                        const char *what) {
  SUMPY_REQUIRE(nlevels >= 1 && nlevels <= LIST_MAXLEV && level_rscale,
                "%s: 1..%d level scales needed, got %d", what, LIST_MAXLEV, nlevels);
  for (int l = 0; l < LIST_MAXLEV; ++l) ls.r[l] = level_rscale[l < nlevels ? l : nlevels - 1];
  return SUMPYHIP_OK;
}

extern "C" int sumpyhip_m2p(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                            const int *target_boxes, const int *source_box_starts,
                            const int *source_box_lists, const int *box_target_starts,
                            const int *box_target_counts_nonchild,
                            const double *box_centers, int nboxes, const double *tx,
                            const double *ty, const double *tz, double rscale,
                            double scale, const double *src_expansions, int src_base_ibox,
                            double *pot, double *gx, double *gy, double *gz,
                            void *stream) {
  LevelScales ls{};
  return m2p_impl(kernel_id, expn, ntgt_boxes, target_boxes, source_box_starts,
                  source_box_lists, box_target_starts, box_target_counts_nonchild, box_centers,
                  nboxes, tx, ty, tz, rscale, scale, src_expansions, src_base_ibox, pot, gx, gy,
                  gz, nullptr, ls, 0, stream);
}

extern "C" int sumpyhip_m2p_levels(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                                   const int *target_boxes, const int *source_box_starts,
                                   const int *source_box_lists, const int *box_target_starts,
                                   const int *box_target_counts_nonchild,
                                   const double *box_centers, int nboxes, const double *tx,
                                   const double *ty, const double *tz, const int *box_levels,
                                   int nlevels, const double *level_rscale, double scale,
                                   const double *src_expansions, double *pot, double *gx,
                                   double *gy, double *gz, int precision, void *stream) {
  SUMPY_REQUIRE(box_levels, "m2p_levels: box_levels needed");
  LevelScales ls;
  int rc = level_scales(nlevels, level_rscale, ls, "m2p_levels");
  if (rc) return rc;
  return m2p_impl(kernel_id, expn, ntgt_boxes, target_boxes, source_box_starts,
                  source_box_lists, box_target_starts, box_target_counts_nonchild, box_centers,
                  nboxes, tx, ty, tz, level_rscale[0], scale, src_expansions, 0, pot, gx, gy, gz,
                  box_levels, ls, precision, stream);
}

extern "C" int sumpyhip_p2l(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                            const int *target_boxes, const int *source_box_starts,
                            const int *source_box_lists, const int *box_source_starts,
                            const int *box_source_counts_nonchild,
                            const double *box_centers, int nboxes, const double *sx,
                            const double *sy, const double *sz, const double *w0,
                            const double *wx, const double *wy, const double *wz,
                            double rscale, double *tgt_expansions, int tgt_base_ibox,
                            void *stream) {
  LevelScales ls{};
  return p2l_impl(kernel_id, expn, ntgt_boxes, target_boxes, source_box_starts,
                  source_box_lists, box_source_starts, box_source_counts_nonchild, box_centers,
                  nboxes, sx, sy, sz, w0, wx, wy, wz, rscale, tgt_expansions, tgt_base_ibox,
                  nullptr, ls, 0, stream);
}

extern "C" int sumpyhip_p2l_levels(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                                   const int *target_boxes, const int *source_box_starts,
                                   const int *source_box_lists, const int *box_source_starts,
                                   const int *box_source_counts_nonchild,
                                   const double *box_centers, int nboxes, const double *sx,
                                   const double *sy, const double *sz, const double *w0,
                                   const double *wx, const double *wy, const double *wz,
                                   const int *box_levels, int nlevels,
                                   const double *level_rscale, double *tgt_expansions,
                                   int precision, void *stream) {
  SUMPY_REQUIRE(box_levels, "p2l_levels: box_levels needed");
  LevelScales ls;
  int rc = level_scales(nlevels, level_rscale, ls, "p2l_levels");
  if (rc) return rc;
  return p2l_impl(kernel_id, expn, ntgt_boxes, target_boxes, source_box_starts,
                  source_box_lists, box_source_starts, box_source_counts_nonchild, box_centers,
                  nboxes, sx, sy, sz, w0, wx, wy, wz, level_rscale[0], tgt_expansions, 0,
                  box_levels, ls, precision, stream);
}

extern "C" int sumpyhip_p2m(int kernel_id, const sumpyhip_expn *expn, int nsrc_boxes,
                            const int *source_boxes, const int *box_source_starts,
                            const int *box_source_counts_nonchild,
                            const double *box_centers, int nboxes, const double *sx,
                            const double *sy, const double *sz, const double *w0,
                            const double *wx, const double *wy, const double *wz,
                            double rscale, double *tgt_expansions, int tgt_base_ibox,
                            void *stream) {
  LevelScales ls{};
  return p2m_impl(kernel_id, expn, nsrc_boxes, source_boxes, box_source_starts,
                  box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz, w0, wx, wy, wz,
                  rscale, tgt_expansions, tgt_base_ibox, nullptr, ls, stream);
}

extern "C" int sumpyhip_p2m_levels(int kernel_id, const sumpyhip_expn *expn, int nsrc_boxes,
                                   const int *source_boxes, const int *box_source_starts,
                                   const int *box_source_counts_nonchild,
                                   const double *box_centers, int nboxes, const double *sx,
                                   const double *sy, const double *sz, const double *w0,
                                   const double *wx, const double *wy, const double *wz,
                                   const int *box_levels, int nlevels,
                                   const double *level_rscale, double *tgt_expansions,
                                   int precision, void *stream) {
  SUMPY_REQUIRE(kernel_id == 3, "p2m_levels: the Laplace 3D kernel only (rscale-free tables)");
  SUMPY_REQUIRE(box_levels, "p2m_levels: box_levels needed");
  SUMPY_REQUIRE(precision == 0 || precision == 1, "p2m_levels: precision 0 (FP64) or 1 (FP32)");
  LevelScales ls;
  int rc = level_scales(nlevels, level_rscale, ls, "p2m_levels");
  if (rc) return rc;
  return p2m_impl(kernel_id, expn, nsrc_boxes, source_boxes, box_source_starts,
                  box_source_counts_nonchild, box_centers, nboxes, sx, sy, sz, w0, wx, wy, wz,
                  level_rscale[0], tgt_expansions, 0, box_levels, ls, stream, precision);
}

extern "C" int sumpyhip_l2p(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                            const int *target_boxes, const int *box_target_starts,
                            const int *box_target_counts_nonchild,
                            const double *box_centers, int nboxes, const double *tx,
                            const double *ty, const double *tz, double rscale,
                            double scale, const double *src_expansions, int src_base_ibox,
                            double *pot, double *gx, double *gy, double *gz,
                            void *stream) {
  LevelScales ls{};
  return l2p_impl(kernel_id, expn, ntgt_boxes, target_boxes, box_target_starts,
                  box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz, rscale, scale,
                  src_expansions, src_base_ibox, pot, gx, gy, gz, nullptr, ls, 0, stream);
}

extern "C" int sumpyhip_l2p_levels(int kernel_id, const sumpyhip_expn *expn, int ntgt_boxes,
                                   const int *target_boxes, const int *box_target_starts,
                                   const int *box_target_counts_nonchild,
                                   const double *box_centers, int nboxes, const double *tx,
                                   const double *ty, const double *tz, const int *box_levels,
                                   int nlevels, const double *level_rscale, double scale,
                                   const double *src_expansions, double *pot, double *gx,
                                   double *gy, double *gz, int precision, void *stream) {
  SUMPY_REQUIRE(kernel_id == 3, "l2p_levels: the Laplace 3D kernel only (rscale-free tables)");
  SUMPY_REQUIRE(box_levels, "l2p_levels: box_levels needed");
  LevelScales ls;
  int rc = level_scales(nlevels, level_rscale, ls, "l2p_levels");
  if (rc) return rc;
  return l2p_impl(kernel_id, expn, ntgt_boxes, target_boxes, box_target_starts,
                  box_target_counts_nonchild, box_centers, nboxes, tx, ty, tz, level_rscale[0],
                  scale, src_expansions, 0, pot, gx, gy, gz, box_levels, ls, precision, stream);
}
