// Multipole-to-local translation through translation classes.
//
// Reference pipeline (sumpy/fmm.py:723-920): M2LGenerateTranslationClassesDependentData
// (e2e.py:565-669) -> M2LPreprocessMultipole (e2e.py:676-760) ->
// M2LUsingTranslationClassesDependentData (e2e.py:310-558) -> M2LPostprocessLocal
// (e2e.py:767-859); by default with FFTs over a circulant embedding
// (m2l.py:913-1001).
//
// Here the class data keep the reference's non-FFT circulant layout
// (m2l.py:403-537): data_c[off(m)] = rho^|m| d^m G(v_c), off() row-major in
// the box [0, max_mi], so that the dense class matrix is
//   A_c[t, s] = data_c[off(t) + off(s)]          (off is linear in m).
// The translation of a level is a class-batched dense contraction on MFMA:
// each block owns a tile of 64 target boxes (grouped on the host by their
// octant within the parent, so that a tile shares its class set), and for every
// class slot of the tile it stages data_c and the 64 source multipoles in LDS
// and accumulates Y(nl x 64) += A_c (nl x nm) X (nm x 64) with
// v_mfma_f64_16x16x4 (plain form) or v_mfma_f32_16x16x4 (FP32 form), A
// fragments gathered straight from data_c in LDS.  Roofline: MFMA.
#include <algorithm>
#include <mutex>

#include "common.h"

namespace sumpyhip {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// {{{ scaled derivatives of 1/r over a row-major box, degree by degree

// D[(mx*M1 + my)*M2 + mz] = rho^|m| d^m (1/r)(v) for |m| <= N; zero outside.
// Recurrence of sumpy/derivative_taker.py:227-259 (d = first axis with
// m_d >= 2); max(m) <= 1 by direct differentiation.  All threads of the block
// participate; ends with a barrier.
__device__ void fill_derivs(double *D, int N, int M0, int M1, int M2, const double v[3],
                            double rscale) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ncirc = M0*M1*M2;
  for (int i = tid; i < ncirc; i += nt) D[i] = 0;
  const double xs[3] = {v[0]/rscale, v[1]/rscale, v[2]/rscale};
  const double r2s = xs[0]*xs[0] + xs[1]*xs[1] + xs[2]*xs[2];
  const double ri = rsqrt(r2s);
  const double rinv = ri/rscale;
  const double ri2 = ri*ri;
  __syncthreads();
  for (int n = 0; n <= N; ++n) {
    for (int idx = tid; idx < M0*M1; idx += nt) {
      const int mx = idx/M1, my = idx % M1, mz = n - mx - my;
      if (mz < 0 || mz >= M2) continue;
      const int mi[3] = {mx, my, mz};
      double val;
      if (mx <= 1 && my <= 1 && mz <= 1) {
        if (n == 0) val = rinv;
        else if (n == 1) val = -xs[mx ? 0 : (my ? 1 : 2)]*ri2*rinv;
        else if (n == 2) {
          int a = mx ? 0 : 1, b = mz ? 2 : 1;
          val = 3.0*xs[a]*xs[b]*ri2*ri2*rinv;
        } else {
          val = -15.0*xs[0]*xs[1]*xs[2]*ri2*ri2*ri2*rinv;
        }
      } else {
        const int d = mx >= 2 ? 0 : (my >= 2 ? 1 : 2);
        double acc = 0;
        for (int i = 0; i < 3; ++i) {
          const int nn = mi[i];
          if (nn == 0) continue;
          int m1[3] = {mi[0], mi[1], mi[2]}; m1[i] -= 1;
          const double v1 = D[(m1[0]*M1 + m1[1])*M2 + m1[2]];
          double v2 = 0;
          if (nn >= 2) {
            int m2[3] = {mi[0], mi[1], mi[2]}; m2[i] -= 2;
            v2 = D[(m2[0]*M1 + m2[1])*M2 + m2[2]];
          }
          if (i == d) acc -= (2*nn - 1)*xs[i]*v1 + double((nn - 1)*(nn - 1))*v2;
          else acc -= 2*nn*xs[i]*v1 + double(nn*(nn - 1))*v2;
        }
        val = acc*ri2;
      }
      D[(mx*M1 + my)*M2 + mz] = val;
    }
    __threadfence_block();
    __syncthreads();
  }
}

// }}}

__global__ __launch_bounds__(256) void m2l_class_data_kernel(
    int N, int M0, int M1, int M2, int nclasses, const double *vecs, int nvectors,
    int class_start, double rscale, double *out) {
  const size_t ncirc = (size_t)M0*M1*M2;
  for (int c = blockIdx.x; c < nclasses; c += gridDim.x) {
    const double v[3] = {vecs[class_start + c], vecs[nvectors + class_start + c],
                         vecs[2*(size_t)nvectors + class_start + c]};
    fill_derivs(out + c*ncirc, N, M0, M1, M2, v, rscale);
  }
}

// packed row entry k holds stored coefficient ksrc[k] (-1: zero; ksrc null:
// entry k holds coefficient k)
__device__ __forceinline__ int k_source_of(const int *ksrc, int k, int nstored) {
  return ksrc ? ksrc[k] : (k < nstored ? k : -1);
}

template <typename T>
__global__ void m2l_preprocess_kernel(int nboxes, int nstored, const double *src, int kpad,
                                      const int *ksrc, T *packed) {
  const size_t n = (size_t)nboxes*kpad;
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x*blockDim.x) {
    const size_t b = i/kpad;
    const int k = k_source_of(ksrc, i % kpad, nstored);
    packed[i] = k >= 0 ? T(src[b*nstored + k]) : T(0);
  }
}

constexpr int MAXDEG = 64;  // ratio^d table of the postprocessing (d = |mi| <= 2 order)

// locals[b, i] = work[b, pos[i]] ratio^deg[i]  (pos null: pos[i] = i)
template <typename T>
__global__ __launch_bounds__(256) void m2l_postprocess_kernel(
    int nboxes, int nl, int npad, const T *work, const int *deg, const int *pos, double ratio,
    double *out) {
  __shared__ double pw[MAXDEG];
  if (threadIdx.x < MAXDEG) pw[threadIdx.x] = pow(ratio, (double)threadIdx.x);
  __syncthreads();
  const size_t n = (size_t)nboxes*nl;
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x*blockDim.x) {
    const size_t b = i/nl;
    const int k = i % nl;
    out[i] = double(work[b*npad + (pos ? pos[k] : k)])*pw[deg[k]];
  }
}

// {{{ the MFMA contraction

constexpr int TW = 64;       // targets per tile (the widest tile; 32 is the other width)
constexpr int MAXRBW = 4;    // row blocks of 16 per wave (npad <= 256)

template <typename T>
struct Mfma;
// Operand layouts (checked on gfx950): A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k; accumulator register r of lane l holds column l % 16 and row
// (l/16) + 4r for the FP64 instruction, 4(l/16) + r for the FP32 one.
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4*r; }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4*(lane >> 4) + r; }
};

// One tree level of a (possibly multi-level) launch.
struct LevelArgs {
  int ntiles, nsplit;   // slot ranges of a tile split over nsplit blocks
  int item_end;         // cumulative (tile, split) items up to this level
  int sbase, tbase;
  const int *tile_targets, *tile_slot_starts, *slot_class, *slot_sources;
  const void *class_data;     // FP64 values, or packed (hi | lo) words at precision 2
  const void *packed;
  void *work;
  void *partial;        // (nsplit - 1) x ntiles*64 x npad sums of splits 1..
};

constexpr int M2L_MAXLEV = 32;

struct ApplyArgs {
  int nlev;
  int nl, nm, npad, kpad, ncirc;
  int tw;               // targets per tile: 64 or 32 (every level of a launch)
  const int *off_t, *off_s;
  LevelArgs lv[M2L_MAXLEV];
};

int set_lds_limit(const void *fn, size_t bytes);

// 16-byte vector type of each compute precision (global loads of X)
template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double2 type;
  static constexpr int N = 2;
};
template <>
struct Vec16<float> {
  typedef float4 type;
  static constexpr int N = 4;
};

// Per-instance limits: RBW row blocks of 16 per wave (npad <= 64 RBW); the
// largest PDE-compressed order that fits sets the class-data size.
template <int RBW>
struct Limits {
  static constexpr int PMAX = RBW == 1 ? 7 : RBW == 2 ? 10 : RBW == 3 ? 12 : 15;
  static constexpr int MAXK = 64*RBW;
  static constexpr int MAXCIRC = 3*(2*PMAX + 1)*(2*PMAX + 1);
  static constexpr int DPT = (MAXCIRC + 255)/256;   // class-data values per thread
};

template <typename T, int RBW, int NCB>
struct Apply {
  static constexpr int VN = Vec16<T>::N;
  static constexpr int TWL = 16*NCB;  // targets per tile (column blocks of 16)
  // padded LDS row of X (FP64: 4 doubles, so that two blocks fit in one CU)
  static constexpr int XS = TWL + (sizeof(T) == 4 ? 16 : 4);
  static constexpr int XPT = (Limits<RBW>::MAXK/VN*TW + 255)/256;  // X vectors per thread
  typedef typename Vec16<T>::type vec_t;
};

// One block per (tile, split) item (4 waves); split j of a tile covers slots
// [n j / nsplit, n (j + 1) / nsplit) of its n slots and writes work (j = 0) or
// partial[j - 1] (the sums are added by m2l_reduce_splits).  Consecutive block
// ids go to different XCDs, so items are mapped XCD-major: an XCD works on a
// contiguous range of tiles, which the schedule orders spatially, and its L2
// holds their source multipoles.  Per class slot:
//   stage   : data_c (ncirc) and X (kpad x 64, LDS row stride XS) in LDS;
//   compute : wave w owns row blocks w, w+4, ... and all four 16-column
//             blocks; per 4-deep k step it reads 4 B fragments and gathers
//             one A fragment per row block from data_c.
// The next slot's data are loaded into registers before the MFMA loop of the
// current slot, so global-memory latency overlaps the contraction; LDS is
// rewritten between two barriers.
template <typename T, int RBW, int NCB>
__global__ __launch_bounds__(256) void m2l_apply_kernel(ApplyArgs a) {
  typedef typename Mfma<T>::acc_t acc_t;
  typedef Apply<T, RBW, NCB> C;
  constexpr int TWL = C::TWL;
  typedef typename C::vec_t vec_t;
  constexpr int DPT = Limits<RBW>::DPT;
  extern __shared__ unsigned char lds_raw[];
  T *sD = reinterpret_cast<T *>(lds_raw);                       // [ncirc]
  T *sX = sD + ((a.ncirc + 3) & ~3);                            // [kpad][XS]
  int *sOff = reinterpret_cast<int *>(sX + (size_t)a.kpad*C::XS);  // [kpad]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kq = a.kpad/C::VN;          // vectors per source row
  // (tile, split) item of this block and its level.  Consecutive block ids
  // go to different XCDs: XCD x takes the x-th eighth of every level's items
  // (a contiguous, spatially ordered range), levels in order.
  const int xcd = blockIdx.x & 7;
  int r = blockIdx.x >> 3, lev = 0, li = -1;
  for (; lev < a.nlev; ++lev) {
    const int n = a.lv[lev].item_end - (lev ? a.lv[lev - 1].item_end : 0);
    const int per = (n + 7)/8, lo = min(xcd*per, n), cnt = min(lo + per, n) - lo;
    if (r < cnt) { li = lo + r; break; }
    r -= cnt;
  }
  if (li < 0) return;
  const LevelArgs L = a.lv[lev];
  const T *packed = reinterpret_cast<const T *>(L.packed);
  T *work = reinterpret_cast<T *>(L.work);
  // Padding rows (>= nl) and columns (>= nm) read an arbitrary valid entry of
  // data_c: the matching X rows are zero (preprocessing pads with zeros) and
  // the padding output rows are never read back.
  for (int k = tid; k < a.kpad; k += 256) sOff[k] = k < a.nm ? a.off_s[k] : 0;
  int offT[RBW];
  for (int r = 0; r < RBW; ++r) {
    const int row = (wave + 4*r)*16 + (lane & 15);
    offT[r] = (row < a.nl) ? a.off_t[row] : 0;
  }

  double dreg[DPT];
  vec_t xreg[C::XPT];
  // global -> registers for one slot; every load is unconditional (clamped
  // address, then a select) so all of them are in flight together
  auto fetch = [&](int slot) {
    const double *D =
        static_cast<const double *>(L.class_data) + (size_t)L.slot_class[slot]*a.ncirc;
    for (int i = 0; i < DPT; ++i) {
      const int e = tid + 256*i;
      dreg[i] = D[e < a.ncirc ? e : 0];
    }
    const int sb = lane < TWL ? L.slot_sources[(size_t)slot*TWL + lane] : -1;
    const T *row = packed + (size_t)((sb >= 0 ? sb : L.sbase) - L.sbase)*a.kpad;
    for (int i = 0; i < C::XPT; ++i) {
      const int q = wave + 4*i;
      xreg[i] = *reinterpret_cast<const vec_t *>(row + (q < kq ? q : 0)*C::VN);
    }
    if (sb < 0)
      for (int i = 0; i < C::XPT; ++i) xreg[i] = vec_t{};
  };
  // registers -> LDS
  auto stage = [&]() {
    for (int i = 0; i < DPT; ++i) {
      const int e = tid + 256*i;
      if (e < a.ncirc) sD[e] = T(dreg[i]);
    }
    for (int i = 0; i < C::XPT; ++i) {
      const int q = wave + 4*i;
      if (q < kq && lane < TWL) {
        const T *xv = reinterpret_cast<const T *>(&xreg[i]);
        for (int j = 0; j < C::VN; ++j) sX[(q*C::VN + j)*C::XS + lane] = xv[j];
      }
    }
  };

  const T *xlane = sX + (lane >> 4)*C::XS + (lane & 15);
  const int *olane = sOff + (lane >> 4);
  {
    const int tile = li/L.nsplit, split = li - tile*L.nsplit;
    acc_t acc[RBW][NCB];
    for (int r = 0; r < RBW; ++r)
      for (int cb = 0; cb < NCB; ++cb) acc[r][cb] = acc_t{0, 0, 0, 0};
    const int t0 = L.tile_slot_starts[tile], nsl = L.tile_slot_starts[tile + 1] - t0;
    const int s0 = t0 + (int)((long long)nsl*split/L.nsplit);
    const int s1 = t0 + (int)((long long)nsl*(split + 1)/L.nsplit);
    if (s0 < s1) {
      fetch(s0);
      __syncthreads();
      stage();
      __syncthreads();
    }
    for (int slot = s0; slot < s1; ++slot) {
      if (slot + 1 < s1) fetch(slot + 1);
      int os_next = olane[0];
      #pragma unroll 2
      for (int k0 = 0; k0 < a.kpad; k0 += 4) {
        // the column offset of the next k step is loaded one step ahead so
        // that the A gather does not wait on a dependent LDS read
        const int os = os_next;
        os_next = olane[k0 + 4 < a.kpad ? k0 + 4 : 0];
        const T *xr = xlane + k0*C::XS;
        T bfr[NCB];
        #pragma unroll
        for (int cb = 0; cb < NCB; ++cb) bfr[cb] = xr[cb*16];
        #pragma unroll
        for (int r = 0; r < RBW; ++r) {
          const T af = sD[offT[r] + os];
          #pragma unroll
          for (int cb = 0; cb < NCB; ++cb) acc[r][cb] = Mfma<T>::mma(af, bfr[cb], acc[r][cb]);
        }
      }
      __syncthreads();
      if (slot + 1 < s1) {
        stage();
        __syncthreads();
      }
    }
    // epilogue: each lane holds four rows of one column of every block
    const int nrb = a.npad >> 4;
    for (int r = 0; r < RBW; ++r) {
      const int rb = wave + 4*r;
      if (rb >= nrb) break;
      for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb*16 + (lane & 15);
        const int tb = L.tile_targets[(size_t)tile*TWL + col];
        if (tb < 0) continue;
        T *dst = split == 0
            ? work + (size_t)(tb - L.tbase)*a.npad
            : reinterpret_cast<T *>(L.partial) +
                  ((size_t)(split - 1)*L.ntiles*TWL + (size_t)tile*TWL + col)*a.npad;
        for (int i = 0; i < 4; ++i) dst[rb*16 + Mfma<T>::row(lane, i)] = acc[r][cb][i];
      }
    }
  }
}

// work[target] += sum over splits 1.. of partial[split - 1][tile position],
// for every level with nsplit > 1 (blockIdx.y = level)
template <typename T>
__global__ __launch_bounds__(256) void m2l_reduce_splits(ApplyArgs a) {
  const LevelArgs &L = a.lv[blockIdx.y];
  if (L.nsplit < 2) return;
  const int npad = a.npad;
  const T *partial = reinterpret_cast<const T *>(L.partial);
  T *work = reinterpret_cast<T *>(L.work);
  const size_t npos = (size_t)L.ntiles*a.tw, n = npos*npad;
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x*blockDim.x) {
    const size_t pos = i/npad;
    const int tb = L.tile_targets[pos];
    if (tb < 0) continue;
    T v = 0;
    for (int j = 0; j < L.nsplit - 1; ++j) v += partial[(size_t)j*n + i];
    work[(size_t)(tb - L.tbase)*npad + (i - pos*npad)] += v;
  }
}

// {{{ split-BF16 contraction (x = hi + lo in bf16; hi*hi + hi*lo + lo*hi, FP32 sums)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

constexpr int KB = 32;        // k per v_mfma_f32_16x16x32_bf16
constexpr int XB = 128 + 8;   // LDS row of one target column (bf16 entries, padded)

__device__ __forceinline__ void split_bf16(float v, __bf16 &hi, __bf16 &lo) {
  hi = static_cast<__bf16>(v);
  lo = static_cast<__bf16>(v - static_cast<float>(hi));
}

__device__ __forceinline__ unsigned bf16_bits(__bf16 h) {
  return (unsigned)__builtin_bit_cast(unsigned short, h);
}

// class data (circulant shape M0 x M1 x M2) -> packed words hi | lo << 16
// (bf16 bit patterns) at mx S0 + my S1 + mz, zero in the gaps
__global__ void m2l_pack_bf16_class_data(int nclasses, int M0, int M1, int M2, int S0, int S1,
                                         int ncirc_p, const double *data, unsigned *out) {
  const size_t n = (size_t)nclasses*ncirc_p, ncirc = (size_t)M0*M1*M2;
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x*blockDim.x) {
    const size_t c = i/ncirc_p;
    const int w = i - c*ncirc_p;
    const int mx = w/S0, r = w - mx*S0, my = r/S1, mz = r - my*S1;
    double v = 0.0;
    if (mx < M0 && my < M1 && mz < M2) v = data[c*ncirc + ((size_t)mx*M1 + my)*M2 + mz];
    __bf16 h, l;
    split_bf16(float(v), h, l);
    out[i] = bf16_bits(h) | (bf16_bits(l) << 16);
  }
}

// stored multipoles -> rows of kpad hi then kpad lo bf16 (entry k holds
// coefficient ksrc[k], zero when -1)
__global__ void m2l_preprocess_bf16_kernel(int nboxes, int nstored, const double *src, int kpad,
                                           const int *ksrc, unsigned short *packed) {
  const size_t n = (size_t)nboxes*kpad;
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x*blockDim.x) {
    const size_t b = i/kpad;
    const int k = i % kpad, ks = k_source_of(ksrc, k, nstored);
    __bf16 h, l;
    split_bf16(ks >= 0 ? float(src[b*nstored + ks]) : 0.f, h, l);
    packed[b*2*kpad + k] = (unsigned short)bf16_bits(h);
    packed[b*2*kpad + kpad + k] = (unsigned short)bf16_bits(l);
  }
}

// Same block/item structure as m2l_apply_kernel.  Inputs come pre-split:
// class data as packed (hi | lo) words, multipole rows as kpad hi then kpad lo
// bf16, so staging is a plain copy.  The next slot's class words and columns
// are fetched into registers while a slot is contracted; per slot they are
// staged in LDS (the tile's 64 hi / lo columns k contiguous).  Every 32-deep
// k step of a wave gathers 8 class words per A row, regroups their halves
// with byte permutes and issues 3 bf16 MFMAs per (row block, column block).
// Within a wave the LDS reads run ahead of the MFMAs that need them (LDS
// returns in order): the B fragments of the first column block, then the A
// words of the next k step, then per column block the B fragments of the
// next one; the kernel is held to 3 waves per SIMD (its register budget).
// Operand layout of v_mfma_f32_16x16x32_bf16: A row = lane % 16, B column =
// lane % 16, the 8 k of a lane are 8 (lane / 16) + e in both; accumulator
// register r = row 4 (lane / 16) + r.
constexpr int BF16_MAXCIRC = 2560;  // packed class words per slot (LDS)

template <int RBW, int NCB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void m2l_apply_bf16_kernel(ApplyArgs a) {
  constexpr int TWL = 16*NCB;  // targets per tile
  constexpr int DPT = (BF16_MAXCIRC + 255)/256;
  extern __shared__ unsigned char lds_raw[];
  unsigned *sD = reinterpret_cast<unsigned *>(lds_raw);                  // [ncirc]
  __bf16 *sXh = reinterpret_cast<__bf16 *>(sD + ((a.ncirc + 3) & ~3));  // [64][XB]
  __bf16 *sXl = sXh + TWL*XB;                                            // [TWL][XB]
  int *sOff = reinterpret_cast<int *>(sXl + TWL*XB);                     // [kpad]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kq = a.kpad/8;  // 8-wide k chunks per row
  const int xcd = blockIdx.x & 7;
  int r = blockIdx.x >> 3, lev = 0, li = -1;
  for (; lev < a.nlev; ++lev) {
    const int n = a.lv[lev].item_end - (lev ? a.lv[lev - 1].item_end : 0);
    const int per = (n + 7)/8, lo = min(xcd*per, n), cnt = min(lo + per, n) - lo;
    if (r < cnt) { li = lo + r; break; }
    r -= cnt;
  }
  if (li < 0) return;
  const LevelArgs L = a.lv[lev];
  const unsigned short *packed = reinterpret_cast<const unsigned short *>(L.packed);
  const unsigned *cwords = reinterpret_cast<const unsigned *>(L.class_data);
  float *work = reinterpret_cast<float *>(L.work);
  for (int k = tid; k < a.kpad; k += 256) sOff[k] = k < a.nm ? a.off_s[k] : 0;
  int offT[RBW];
  for (int q = 0; q < RBW; ++q) {
    const int row = (wave + 4*q)*16 + (lane & 15);
    offT[q] = (row < a.nl) ? a.off_t[row] : 0;
  }
  // one slot's multipole columns: for column `lane`, the hi and lo 8-chunks
  // m = wave + 4 i (i < 4)
  struct XRegs {
    u4 xh[4], xl[4];
  };
  // slot range of this block (computed here so that loads can be clamped to
  // it: every fetch is unconditional, which keeps the memory-counter waits
  // exact)
  const int tile = li/L.nsplit, split = li - tile*L.nsplit;
  const int t0 = L.tile_slot_starts[tile], nsl = L.tile_slot_starts[tile + 1] - t0;
  const int s0 = t0 + (int)((long long)nsl*split/L.nsplit);
  const int s1 = t0 + (int)((long long)nsl*(split + 1)/L.nsplit);
  // source box of this lane's column in a slot (loaded one fetch ahead)
  auto meta = [&](int slot) {
    return lane < TWL ? L.slot_sources[(size_t)min(slot, s1 - 1)*TWL + lane] : -1;
  };
  auto fetch_d = [&](int slot, unsigned (&d)[DPT]) {
    const unsigned *D = cwords + (size_t)L.slot_class[min(slot, s1 - 1)]*a.ncirc;
    for (int i = 0; i < DPT; ++i) {
      const int e = tid + 256*i;
      d[i] = D[e < a.ncirc ? e : 0];
    }
  };
  auto fetch_x = [&](int sb, XRegs &g) {
    const unsigned short *row = packed + (size_t)((sb >= 0 ? sb : L.sbase) - L.sbase)*2*a.kpad;
    for (int i = 0; i < 4; ++i) {
      const int m = wave + 4*i, mm = m < kq ? m : 0;
      g.xh[i] = *reinterpret_cast<const u4 *>(row + mm*8);
      g.xl[i] = *reinterpret_cast<const u4 *>(row + a.kpad + mm*8);
    }
    if (sb < 0)
      for (int i = 0; i < 4; ++i) g.xh[i] = g.xl[i] = u4{0, 0, 0, 0};
  };
  auto stage = [&](const unsigned (&d)[DPT], const XRegs &g) {
    for (int i = 0; i < DPT; ++i) {
      const int e = tid + 256*i;
      if (e < a.ncirc) sD[e] = d[i];
    }
    for (int i = 0; i < 4; ++i) {
      const int m = wave + 4*i;
      if (m < kq && lane < TWL) {
        *reinterpret_cast<u4 *>(sXh + lane*XB + m*8) = g.xh[i];
        *reinterpret_cast<u4 *>(sXl + lane*XB + m*8) = g.xl[i];
      }
    }
  };
  typedef float v4f_t __attribute__((ext_vector_type(4)));
  v4f_t acc[RBW][NCB];
  for (int q = 0; q < RBW; ++q)
    for (int cb = 0; cb < NCB; ++cb) acc[q][cb] = v4f_t{0, 0, 0, 0};
  const int kg = (lane >> 4)*8;  // first k of this lane inside a 32-deep step
  const int nks = a.kpad/KB;
  auto gather = [&](int k0, unsigned (&w)[RBW][8]) {
    int os[8];
    #pragma unroll
    for (int e = 0; e < 8; ++e) os[e] = sOff[k0 + kg + e];
    #pragma unroll
    for (int q = 0; q < RBW; ++q)
      #pragma unroll
      for (int e = 0; e < 8; ++e) w[q][e] = sD[offT[q] + os[e]];
  };
  auto to_frag = [&](const unsigned (&w)[RBW][8], bf16x8 (&ah)[RBW], bf16x8 (&al)[RBW]) {
    #pragma unroll
    for (int q = 0; q < RBW; ++q) {
      u4 hw, lw;
      #pragma unroll
      for (int e = 0; e < 4; ++e) {
        hw[e] = __builtin_amdgcn_perm(w[q][2*e + 1], w[q][2*e], 0x05040100u);
        lw[e] = __builtin_amdgcn_perm(w[q][2*e + 1], w[q][2*e], 0x07060302u);
      }
      ah[q] = __builtin_bit_cast(bf16x8, hw);
      al[q] = __builtin_bit_cast(bf16x8, lw);
    }
  };
  auto contract = [&]() {
    unsigned wn[RBW][8];
    bf16x8 ah[RBW], al[RBW];
    gather(0, wn);
    to_frag(wn, ah, al);
    #pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      if (ks >= nks) break;
      const int k0 = ks*KB;
      // B of column block 0, then the A words of the next k step, then per
      // column block: B of the next block in flight while its MFMAs run
      bf16x8 bh[2], bl[2];
      bh[0] = *reinterpret_cast<const bf16x8 *>(sXh + (lane & 15)*XB + k0 + kg);
      bl[0] = *reinterpret_cast<const bf16x8 *>(sXl + (lane & 15)*XB + k0 + kg);
      if (ks + 1 < nks) gather(k0 + KB, wn);
      #pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        if (cb + 1 < NCB) {
          const int col = (cb + 1)*16 + (lane & 15);
          bh[(cb + 1) & 1] = *reinterpret_cast<const bf16x8 *>(sXh + col*XB + k0 + kg);
          bl[(cb + 1) & 1] = *reinterpret_cast<const bf16x8 *>(sXl + col*XB + k0 + kg);
        }
        #pragma unroll
        for (int q = 0; q < RBW; ++q) {
          acc[q][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[q], bh[cb & 1], acc[q][cb], 0, 0, 0);
          acc[q][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[q], bl[cb & 1], acc[q][cb], 0, 0, 0);
          acc[q][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[q], bh[cb & 1], acc[q][cb], 0, 0, 0);
        }
      }
      if (ks + 1 < nks) to_frag(wn, ah, al);
    }
  };
  if (s0 < s1) {
    // the class words and columns of the next slot are fetched (clamped, so
    // fetches past the range reload the last slot) while this one runs
    unsigned dn[DPT];
    XRegs ga;
    fetch_d(s0, dn);
    fetch_x(meta(s0), ga);
    stage(dn, ga);
    __syncthreads();
    const int m1 = meta(s0 + 1);
    fetch_d(s0 + 1, dn);
    fetch_x(m1, ga);
    int m2 = meta(s0 + 2);
    for (int slot = s0; slot < s1; ++slot) {
      contract();
      __syncthreads();
      if (slot + 1 >= s1) break;
      stage(dn, ga);
      __syncthreads();
      fetch_d(slot + 2, dn);
      fetch_x(m2, ga);
      m2 = meta(slot + 3);
    }
  }
  const int nrb = a.npad >> 4;
  for (int q = 0; q < RBW; ++q) {
    const int rb = wave + 4*q;
    if (rb >= nrb) break;
    for (int cb = 0; cb < NCB; ++cb) {
      const int col = cb*16 + (lane & 15);
      const int tb = L.tile_targets[(size_t)tile*TWL + col];
      if (tb < 0) continue;
      float *dst = split == 0
          ? work + (size_t)(tb - L.tbase)*a.npad
          : reinterpret_cast<float *>(L.partial) +
                ((size_t)(split - 1)*L.ntiles*TWL + (size_t)tile*TWL + col)*a.npad;
      for (int i = 0; i < 4; ++i) dst[rb*16 + 4*(lane >> 4) + i] = acc[q][cb][i];
    }
  }
}

template <int RBW, int NCB>
int launch_apply_bf16(const ApplyArgs &a, hipStream_t st) {
  SUMPY_REQUIRE(a.kpad % KB == 0 && a.kpad <= 128,
                "m2l_apply (bf16x3): kpad %d must be a multiple of 32 and <= 128", a.kpad);
  const size_t lds = sizeof(unsigned)*((size_t)((a.ncirc + 3) & ~3)) +
                     2*sizeof(__bf16)*(16*NCB)*XB + sizeof(int)*a.kpad;
  SUMPY_REQUIRE(lds <= 160*1024, "m2l_apply: %zu bytes of LDS needed (order too high)", lds);
  int rc = set_lds_limit((const void *)m2l_apply_bf16_kernel<RBW, NCB>, lds);
  if (rc) return rc;
  long long per_xcd = 0;
  for (int l = 0; l < a.nlev; ++l)
    per_xcd += (a.lv[l].item_end - (l ? a.lv[l - 1].item_end : 0) + 7)/8;
  m2l_apply_bf16_kernel<RBW, NCB><<<(int)(8*per_xcd), 256, lds, st>>>(a);
  return check_launch("m2l_apply_bf16");
}

int dispatch_apply_bf16(const ApplyArgs &a, hipStream_t st) {
  const int nrb = a.npad/16;
  int rc;
  const bool narrow = a.tw == 32;
  if (nrb <= 4 && a.ncirc <= BF16_MAXCIRC)
    rc = narrow ? launch_apply_bf16<1, 2>(a, st) : launch_apply_bf16<1, 4>(a, st);
  else if (nrb <= 8 && a.ncirc <= BF16_MAXCIRC)
    rc = narrow ? launch_apply_bf16<2, 2>(a, st) : launch_apply_bf16<2, 4>(a, st);
  else
    return set_error(SUMPYHIP_ENOTIMPL,
                     "m2l_apply (bf16x3): npad %d / ncirc %d exceed the supported sizes",
                     a.npad, a.ncirc);
  if (rc) return rc;
  size_t nmax = 0;
  for (int l = 0; l < a.nlev; ++l)
    if (a.lv[l].nsplit > 1) nmax = std::max(nmax, (size_t)a.lv[l].ntiles*a.tw*a.npad);
  if (nmax == 0) return SUMPYHIP_OK;
  const int rgrid = (int)((nmax + 255)/256 < 8192 ? (nmax + 255)/256 : 8192);
  m2l_reduce_splits<float><<<dim3(rgrid, a.nlev), 256, 0, st>>>(a);
  return check_launch("m2l_reduce_splits");
}

// }}}

template <typename T, int RBW, int NCB>
int launch_apply(const ApplyArgs &a, hipStream_t st) {
  const size_t lds = sizeof(T)*((size_t)((a.ncirc + 3) & ~3) +
                                (size_t)a.kpad*Apply<T, RBW, NCB>::XS) +
                     sizeof(int)*a.kpad;
  SUMPY_REQUIRE(lds <= 160*1024, "m2l_apply: %zu bytes of LDS needed (order too high)", lds);
  int rc = set_lds_limit((const void *)m2l_apply_kernel<T, RBW, NCB>, lds);
  if (rc) return rc;
  long long per_xcd = 0;  // items of XCD 0 (the largest share)
  for (int l = 0; l < a.nlev; ++l)
    per_xcd += (a.lv[l].item_end - (l ? a.lv[l - 1].item_end : 0) + 7)/8;
  const int grid = (int)(8*per_xcd);
  m2l_apply_kernel<T, RBW, NCB><<<grid, 256, lds, st>>>(a);
  rc = check_launch("m2l_apply");
  if (rc) return rc;
  size_t nmax = 0;
  for (int l = 0; l < a.nlev; ++l)
    if (a.lv[l].nsplit > 1) nmax = std::max(nmax, (size_t)a.lv[l].ntiles*a.tw*a.npad);
  if (nmax == 0) return SUMPYHIP_OK;
  const int rgrid = (int)((nmax + 255)/256 < 8192 ? (nmax + 255)/256 : 8192);
  m2l_reduce_splits<T><<<dim3(rgrid, a.nlev), 256, 0, st>>>(a);
  return check_launch("m2l_reduce_splits");
}

template <typename T, int RBW>
int launch_apply_tw(const ApplyArgs &a, hipStream_t st) {
  return a.tw == 32 ? launch_apply<T, RBW, 2>(a, st) : launch_apply<T, RBW, 4>(a, st);
}

template <typename T>
int dispatch_apply(const ApplyArgs &a, hipStream_t st) {
  const int nrb = a.npad/16;
  if (nrb <= 4 && a.kpad <= Limits<1>::MAXK && a.ncirc <= Limits<1>::MAXCIRC)
    return launch_apply_tw<T, 1>(a, st);
  if (nrb <= 8 && a.kpad <= Limits<2>::MAXK && a.ncirc <= Limits<2>::MAXCIRC)
    return launch_apply_tw<T, 2>(a, st);
  if (nrb <= 12 && a.kpad <= Limits<3>::MAXK && a.ncirc <= Limits<3>::MAXCIRC)
    return launch_apply_tw<T, 3>(a, st);
  if (nrb <= 16 && a.kpad <= Limits<4>::MAXK && a.ncirc <= Limits<4>::MAXCIRC)
    return launch_apply_tw<T, 4>(a, st);
  return set_error(SUMPYHIP_ENOTIMPL,
                   "m2l_apply: npad %d / kpad %d / ncirc %d exceed the supported sizes",
                   a.npad, a.kpad, a.ncirc);
}

// {{{ M2L without translation classes: derivatives per pair, LDS matvec

__global__ __launch_bounds__(256) void m2l_csr_kernel(
    int N, int M0, int M1, int M2, int nl, int nm, const int *off_t, const int *off_s,
    const int *deg, int ntgt, const int *target_boxes, const int *starts, const int *lists,
    const double *centers, int nboxes, double src_rscale, double tgt_rscale,
    const double *src, int sbase, double *tgt, int tbase) {
  extern __shared__ double lds[];
  const int ncirc = M0*M1*M2;
  double *D = lds, *Ms = D + ncirc;
  const int tid = threadIdx.x;
  const double ratio = tgt_rscale/src_rscale;
  for (int ib = blockIdx.x; ib < ntgt; ib += gridDim.x) {
    const int tb = target_boxes[ib];
    double acc[4] = {0, 0, 0, 0};
    for (int is = starts[ib]; is < starts[ib + 1]; ++is) {
      const int sb = lists[is];
      __syncthreads();
      for (int k = tid; k < nm; k += 256) Ms[k] = src[(size_t)(sb - sbase)*nm + k];
      const double v[3] = {centers[tb] - centers[sb], centers[nboxes + tb] - centers[nboxes + sb],
                           centers[2*(size_t)nboxes + tb] - centers[2*(size_t)nboxes + sb]};
      fill_derivs(D, N, M0, M1, M2, v, src_rscale);
      for (int q = 0; q < 4; ++q) {
        const int row = tid + 256*q;
        if (row >= nl) break;
        const int ot = off_t[row];
        double s = 0;
        for (int k = 0; k < nm; ++k) s += Ms[k]*D[ot + off_s[k]];
        acc[q] += s;
      }
    }
    for (int q = 0; q < 4; ++q) {
      const int row = tid + 256*q;
      if (row >= nl) break;
      tgt[(size_t)(tb - tbase)*nl + row] = acc[q]*pow(ratio, (double)deg[row]);
    }
  }
}

// }}}

int set_lds_limit(const void *fn, size_t bytes) {
  if (bytes <= 64*1024) return SUMPYHIP_OK;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess)
    return set_error(SUMPYHIP_EHIP, "m2l: cannot raise LDS limit to %zu: %s", bytes,
                     hipGetErrorString(e));
  return SUMPYHIP_OK;
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" int sumpyhip_m2l_class_data(int kernel_id, int order_src, int order_tgt,
                                       const int *circ_shape, int nclasses,
                                       const double *translation_vectors, int nvectors,
                                       int class_start, double src_rscale, double *out,
                                       void *stream) {
  SUMPY_REQUIRE(kernel_id == 3, "m2l_class_data: only the Laplace 3D kernel is implemented");
  SUMPY_REQUIRE(circ_shape != nullptr, "m2l_class_data: null shape");
  if (nclasses == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(class_start >= 0 && class_start + nclasses <= nvectors,
                "m2l_class_data: class range outside the vectors");
  const int grid = nclasses < 1024 ? nclasses : 1024;
  m2l_class_data_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(
      order_src + order_tgt, circ_shape[0], circ_shape[1], circ_shape[2], nclasses,
      translation_vectors, nvectors, class_start, src_rscale, out);
  return check_launch("m2l_class_data");
}

extern "C" int sumpyhip_m2l_preprocess(int precision, int nboxes, int nstored,
                                       const double *src_expansions, int kpad,
                                       const int *k_source, void *packed, void *stream) {
  SUMPY_REQUIRE(kpad >= nstored, "m2l_preprocess: kpad < nstored");
  if (nboxes == 0) return SUMPYHIP_OK;
  const size_t n = (size_t)nboxes*kpad;
  const int grid = (int)((n + 255)/256 < 65536 ? (n + 255)/256 : 65536);
  SUMPY_REQUIRE(precision >= 0 && precision <= 2, "m2l_preprocess: unknown precision");
  if (precision == 0)
    m2l_preprocess_kernel<double><<<grid, 256, 0, (hipStream_t)stream>>>(
        nboxes, nstored, src_expansions, kpad, k_source, (double *)packed);
  else if (precision == 2)
    m2l_preprocess_bf16_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(
        nboxes, nstored, src_expansions, kpad, k_source, (unsigned short *)packed);
  else
    m2l_preprocess_kernel<float><<<grid, 256, 0, (hipStream_t)stream>>>(
        nboxes, nstored, src_expansions, kpad, k_source, (float *)packed);
  return check_launch("m2l_preprocess");
}

extern "C" int sumpyhip_m2l_postprocess(int precision, int nboxes, int nl, int npad,
                                        const void *work, const int *tgt_degree,
                                        const int *row_pos, double ratio,
                                        double *tgt_expansions, void *stream) {
  SUMPY_REQUIRE(npad >= nl, "m2l_postprocess: npad < nl");
  if (nboxes == 0) return SUMPYHIP_OK;
  const size_t n = (size_t)nboxes*nl;
  const int grid = (int)((n + 255)/256 < 65536 ? (n + 255)/256 : 65536);
  SUMPY_REQUIRE(precision >= 0 && precision <= 2, "m2l_postprocess: unknown precision");
  if (precision == 0)
    m2l_postprocess_kernel<double><<<grid, 256, 0, (hipStream_t)stream>>>(
        nboxes, nl, npad, (const double *)work, tgt_degree, row_pos, ratio, tgt_expansions);
  else
    m2l_postprocess_kernel<float><<<grid, 256, 0, (hipStream_t)stream>>>(
        nboxes, nl, npad, (const float *)work, tgt_degree, row_pos, ratio, tgt_expansions);
  return check_launch("m2l_postprocess");
}

namespace sumpyhip {
namespace {

// Fixed device scratch of sumpyhip_m2l_apply_noalloc: the split partial sums
// of one call.  Every GPU the library runs on holds one copy until the
// process ends; concurrent calls on one GPU share it (and race).
constexpr size_t M2L_SCRATCH_BYTES = (size_t)SUMPYHIP_M2L_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[M2L_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void *scratch_ptr[MAX_DEVICES];

// The device address of the current device's copy of scratch (the host-side
// name is only a stand-in), looked up once per device with it current.
int scratch_address(void **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return set_error(SUMPYHIP_EHIP, "m2l_apply_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  SUMPY_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "m2l_apply_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(SUMPYHIP_EHIP, "m2l_apply_noalloc: hipGetSymbolAddress(scratch): %s",
                       hipGetErrorString(e));
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return SUMPYHIP_OK;
}

size_t precision_bytes(int precision) { return precision == 0 ? sizeof(double) : sizeof(float); }

}  // namespace
}  // namespace sumpyhip

extern "C" int sumpyhip_m2l_apply(int precision, int nlevels,
                                  const sumpyhip_m2l_level *levels, int nl, int nm, int npad,
                                  int kpad, const int *off_t, const int *off_s, int ncirc,
                                  void *stream) {
  SUMPY_REQUIRE(precision >= 0 && precision <= 2, "m2l_apply: unknown precision %d", precision);
  SUMPY_REQUIRE(npad % 16 == 0 && npad >= nl, "m2l_apply: npad %d must be a multiple of 16 >= nl",
                npad);
  SUMPY_REQUIRE(kpad % 4 == 0 && kpad >= nm, "m2l_apply: kpad %d must be a multiple of 4 >= nm",
                kpad);
  SUMPY_REQUIRE(nlevels >= 0 && nlevels <= M2L_MAXLEV, "m2l_apply: %d levels (at most %d)",
                nlevels, M2L_MAXLEV);
  ApplyArgs a{};
  a.nl = nl; a.nm = nm; a.npad = npad; a.kpad = kpad; a.ncirc = ncirc;
  a.off_t = off_t; a.off_s = off_s;
  a.tw = 0;
  long long items = 0;
  for (int l = 0; l < nlevels; ++l) {
    const sumpyhip_m2l_level &in = levels[l];
    if (in.ntiles == 0) continue;
    const int tw = in.tile_width ? in.tile_width : TW;
    SUMPY_REQUIRE(tw == 64 || tw == 32, "m2l_apply: tile width %d (64 or 32)", tw);
    SUMPY_REQUIRE(a.tw == 0 || a.tw == tw, "m2l_apply: levels of one launch share the tile width");
    a.tw = tw;
    SUMPY_REQUIRE(in.ntiles > 0 && in.nsplit >= 1, "m2l_apply: bad tiles/splits");
    SUMPY_REQUIRE(in.nsplit == 1 || in.partial != nullptr,
                  "m2l_apply: nsplit > 1 needs a partial buffer");
    items += (long long)in.ntiles*in.nsplit;
    SUMPY_REQUIRE(items < (1ll << 30), "m2l_apply: too many tiles x splits");
    LevelArgs &o = a.lv[a.nlev++];
    o.ntiles = in.ntiles; o.nsplit = in.nsplit; o.item_end = (int)items;
    o.sbase = in.src_base_ibox; o.tbase = in.tgt_base_ibox;
    o.tile_targets = in.tile_targets; o.tile_slot_starts = in.tile_slot_starts;
    o.slot_class = in.slot_class; o.slot_sources = in.slot_sources;
    o.class_data = in.class_data; o.packed = in.packed_src; o.work = in.work;
    o.partial = in.partial;
  }
  if (a.nlev == 0) return SUMPYHIP_OK;
  hipStream_t st = (hipStream_t)stream;
  if (precision == 2) return dispatch_apply_bf16(a, st);
  return precision == 0 ? dispatch_apply<double>(a, st) : dispatch_apply<float>(a, st);
}

extern "C" int sumpyhip_m2l_apply_noalloc(int precision, int nlevels,
                                          const sumpyhip_m2l_level *levels, int nl, int nm,
                                          int npad, int kpad, const int *off_t,
                                          const int *off_s, int ncirc, void *stream) {
  SUMPY_REQUIRE(precision >= 0 && precision <= 2, "m2l_apply_noalloc: unknown precision %d",
                precision);
  SUMPY_REQUIRE(nlevels >= 0 && nlevels <= M2L_MAXLEV, "m2l_apply_noalloc: %d levels (at most %d)",
                nlevels, M2L_MAXLEV);
  SUMPY_REQUIRE(nlevels == 0 || levels != nullptr, "m2l_apply_noalloc: null levels");
  // the split partial sums of every level, carved out of the fixed scratch
  // (256-byte aligned regions); every entry is written before it is read, so
  // nothing needs clearing between calls
  sumpyhip_m2l_level lv[M2L_MAXLEV];
  size_t need = 0;
  for (int l = 0; l < nlevels; ++l) {
    lv[l] = levels[l];
    lv[l].partial = nullptr;
    if (lv[l].ntiles > 0 && lv[l].nsplit > 1) {
      const size_t tw = lv[l].tile_width ? lv[l].tile_width : TW;
      const size_t bytes = (size_t)(lv[l].nsplit - 1)*(size_t)lv[l].ntiles*tw*(size_t)npad*
                           precision_bytes(precision);
      need += (bytes + 255) & ~(size_t)255;
    }
  }
  SUMPY_REQUIRE(need <= M2L_SCRATCH_BYTES,
                "m2l_apply_noalloc: the split partial sums need %zu bytes, the fixed scratch "
                "holds %zu (use fewer splits or sumpyhip_m2l_apply)", need, M2L_SCRATCH_BYTES);
  if (need) {
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    size_t off = 0;
    for (int l = 0; l < nlevels; ++l) {
      if (lv[l].ntiles > 0 && lv[l].nsplit > 1) {
        lv[l].partial = static_cast<unsigned char *>(base) + off;
        const size_t tw = lv[l].tile_width ? lv[l].tile_width : TW;
        const size_t bytes = (size_t)(lv[l].nsplit - 1)*(size_t)lv[l].ntiles*tw*(size_t)npad*
                             precision_bytes(precision);
        off += (bytes + 255) & ~(size_t)255;
      }
    }
  }
  return sumpyhip_m2l_apply(precision, nlevels, lv, nl, nm, npad, kpad, off_t, off_s, ncirc,
                            stream);
}

extern "C" int sumpyhip_m2l_csr(int kernel_id, int order_src, int order_tgt,
                                const int *circ_shape, int nl, int nm, const int *off_t,
                                const int *off_s, const int *tgt_degree, int ntgt_boxes,
                                const int *target_boxes, const int *src_box_starts,
                                const int *src_box_lists, const double *box_centers,
                                int nboxes, double src_rscale, double tgt_rscale,
                                const double *src_expansions, int src_base_ibox,
                                double *tgt_expansions, int tgt_base_ibox, void *stream) {
  SUMPY_REQUIRE(kernel_id == 3, "m2l_csr: only the Laplace 3D kernel is implemented");
  SUMPY_REQUIRE(nl <= 1024, "m2l_csr: too many target coefficients");
  if (ntgt_boxes == 0) return SUMPYHIP_OK;
  const size_t ncirc = (size_t)circ_shape[0]*circ_shape[1]*circ_shape[2];
  const size_t lds = sizeof(double)*(ncirc + nm);
  SUMPY_REQUIRE(lds <= 160*1024, "m2l_csr: %zu bytes of LDS needed (order too high)", lds);
  int rc = set_lds_limit((const void *)m2l_csr_kernel, lds);
  if (rc) return rc;
  const int grid = ntgt_boxes < 8192 ? ntgt_boxes : 8192;
  m2l_csr_kernel<<<grid, 256, lds, (hipStream_t)stream>>>(
      order_src + order_tgt, circ_shape[0], circ_shape[1], circ_shape[2], nl, nm, off_t,
      off_s, tgt_degree, ntgt_boxes, target_boxes, src_box_starts, src_box_lists,
      box_centers, nboxes, src_rscale, tgt_rscale, src_expansions, src_base_ibox,
      tgt_expansions, tgt_base_ibox);
  return check_launch("m2l_csr");
}

extern "C" int sumpyhip_m2l_pack_bf16_class_data(int nclasses, const int *circ_shape, int S0,
                                                 int S1, int ncirc_packed,
                                                 const double *class_data, void *packed,
                                                 void *stream) {
  SUMPY_REQUIRE(nclasses >= 0 && circ_shape, "m2l_pack_bf16_class_data: bad arguments");
  const int M0 = circ_shape[0], M1 = circ_shape[1], M2 = circ_shape[2];
  SUMPY_REQUIRE(M0 > 0 && M1 > 0 && M2 > 0 && S1 >= M2 && S0 >= (M1 - 1)*S1 + M2 &&
                    ncirc_packed >= (M0 - 1)*S0 + (M1 - 1)*S1 + M2,
                "m2l_pack_bf16_class_data: strides (%d, %d) / size %d do not hold shape "
                "(%d, %d, %d)", S0, S1, ncirc_packed, M0, M1, M2);
  if (nclasses == 0) return SUMPYHIP_OK;
  const size_t n = (size_t)nclasses*ncirc_packed;
  const int grid = (int)((n + 255)/256 < 65536 ? (n + 255)/256 : 65536);
  m2l_pack_bf16_class_data<<<grid, 256, 0, (hipStream_t)stream>>>(
      nclasses, M0, M1, M2, S0, S1, ncirc_packed, class_data, (unsigned *)packed);
  return check_launch("m2l_pack_bf16_class_data");
}
