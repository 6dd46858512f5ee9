// Level translations of an octree (M2M and L2L) as dense products on MFMA.
//
// Reference: E2EFromChildren (sumpy/e2e.py:866-978, math in
// sumpy/expansion/multipole.py:159-407) and E2EFromParent (sumpy/e2e.py:985-1083,
// math in sumpy/expansion/local.py:348-493).  Both are linear maps of the
// coefficient vector that depend only on the offset between parent and child
// centre (scaled by rscale) and on the two rscales.  In an octree every child
// of a level sits at one of eight fixed offsets from its parent, so a level's
// M2M is  M_parent += sum_o T_o M_child(o)  and its L2L is
// L_child += T'_o L_parent,  with eight fixed n_out x n_in matrices T_o.  The
// matrices are built once (by the per-box kernels of expansions.hip applied to
// unit vectors) and the translation of a level is a batched product:
//   one block of 8 waves per group of 16 target boxes; per octant present in
//   the group, the 16 input vectors are staged in LDS (B operand, zero padded)
//   and every wave multiplies its row block(s) of T_o (A fragments streamed
//   from L2 in MFMA fragment order, the next chunk in flight) with
//   v_mfma_f64_16x16x4f64; the sums are added to the targets (accumulate).
// Roofline: MFMA (FP64); 2 n_out n_in flops per (target, input) pair.
#include "common.h"

namespace sumpyhip {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GW = 16;     // target boxes per group (one MFMA column block)
constexpr int NWAVE = 8;   // waves per block; wave w owns row blocks w, w + 8, ...
constexpr int KC = 8;      // k steps per prefetched chunk of A fragments
constexpr int BS = GW + 1; // LDS row stride of the staged inputs (doubles)

struct OctArgs {
  int ngroups;
  const int *grp_targets;  // ngroups x 16 (-1: none)
  const int *grp_inputs;   // ngroups x 8 x 16 (-1: none)
  const int *grp_mask;     // ngroups: bit o set when octant o has an input
  int n_out, n_in, nrb, ksteps;
  const double *T;         // 8 x nrb x ksteps x 64 (A-fragment order)
  const double *src;
  int sbase;
  double *tgt;
  int tbase;
};

template <int RBW>
__global__ __launch_bounds__(64*NWAVE) void octant_translate_kernel(OctArgs a) {
  extern __shared__ double sB[];  // [kpad][BS]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kpad = a.ksteps*4;
  const bool active = wave < a.nrb;
  for (int g = blockIdx.x; g < a.ngroups; g += gridDim.x) {
    v4d acc[RBW];
    for (int r = 0; r < RBW; ++r) acc[r] = v4d{0, 0, 0, 0};
    const int mask = a.grp_mask[g];
    for (int o = 0; o < 8; ++o) {
      if (!((mask >> o) & 1)) continue;
      __syncthreads();
      // inputs of octant o, k fastest (coalesced rows), zero padded
      for (int idx = tid; idx < GW*kpad; idx += 64*NWAVE) {
        const int j = idx/kpad, k = idx - j*kpad;
        const int box = a.grp_inputs[((size_t)g*8 + o)*GW + j];
        sB[k*BS + j] = (box >= 0 && k < a.n_in) ? a.src[(size_t)(box - a.sbase)*a.n_in + k] : 0.0;
      }
      __syncthreads();
      if (!active) continue;
      const double *Tw[RBW];
      for (int r = 0; r < RBW; ++r) {
        const int rb = wave + NWAVE*r;
        Tw[r] = a.T + ((size_t)o*a.nrb + (rb < a.nrb ? rb : 0))*a.ksteps*64 + lane;
      }
      double acur[RBW][KC], anxt[RBW][KC];
      for (int r = 0; r < RBW; ++r)
        for (int kk = 0; kk < KC; ++kk)
          acur[r][kk] = Tw[r][(size_t)min(kk, a.ksteps - 1)*64];
      for (int c0 = 0; c0 < a.ksteps; c0 += KC) {
        for (int r = 0; r < RBW; ++r)
          for (int kk = 0; kk < KC; ++kk)
            anxt[r][kk] = Tw[r][(size_t)min(c0 + KC + kk, a.ksteps - 1)*64];
        #pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
          const int ks = c0 + kk;
          if (ks < a.ksteps) {
            const double b = sB[(ks*4 + (lane >> 4))*BS + (lane & 15)];
            #pragma unroll
            for (int r = 0; r < RBW; ++r)
              acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(acur[r][kk], b, acc[r], 0, 0, 0);
          }
        }
        for (int r = 0; r < RBW; ++r)
          for (int kk = 0; kk < KC; ++kk) acur[r][kk] = anxt[r][kk];
      }
    }
    // register i of lane l: row (l/16) + 4i of column l%16
    const int tb = a.grp_targets[(size_t)g*GW + (lane & 15)];
    if (active && tb >= 0) {
      double *dst = a.tgt + (size_t)(tb - a.tbase)*a.n_out;
      for (int r = 0; r < RBW; ++r) {
        const int rb = wave + NWAVE*r;
        if (rb >= a.nrb) break;
        for (int i = 0; i < 4; ++i) {
          const int row = rb*16 + (lane >> 4) + 4*i;
          if (row < a.n_out) dst[row] += acc[r][i];
        }
      }
    }
  }
}

// Split form for n_out <= 128 (one row block per wave) and n_in <= 128: one
// block per (group, octant), so the octants of a group run side by side.  A
// block loads its whole A fragments (one row block per wave) into registers
// while it stages the octant's 16 inputs (zero padded to 128 coefficients),
// then runs 32 MFMA k steps.  With a partial buffer the sums are written per
// octant and added to the targets by octant_reduce in octant order (M2M: a
// parent has inputs in several octants); without one, each target column
// with an input in the block's octant is added to directly (L2L: a child has
// its parent in exactly one octant).
constexpr int KS = 32;                           // k steps (n_in <= 128)
constexpr int SPT = (GW*KS*4 + 64*NWAVE - 1)/(64*NWAVE);  // staged inputs per thread

__global__ __launch_bounds__(64*NWAVE) void octant_translate_split(OctArgs a, const int *items,
                                                                   double *partial) {
  extern __shared__ double sB[];  // [KS*4][BS]
  const int item = items[blockIdx.x];
  const int g = item >> 3, o = item & 7;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool active = wave < a.nrb;
  const int rb = active ? wave : 0;
  const int *inputs = a.grp_inputs + ((size_t)g*8 + o)*GW;
  double A[KS];
  const double *Tw = a.T + ((size_t)o*a.nrb + rb)*a.ksteps*64 + lane;
  #pragma unroll
  for (int ks = 0; ks < KS; ++ks) A[ks] = Tw[(size_t)min(ks, a.ksteps - 1)*64];
  #pragma unroll
  for (int i = 0; i < SPT; ++i) {
    const int idx = tid + 64*NWAVE*i;
    const int j = idx/(KS*4), k = idx - j*(KS*4);
    if (j < GW) {
      const int box = inputs[j];
      sB[k*BS + j] = (box >= 0 && k < a.n_in) ? a.src[(size_t)(box - a.sbase)*a.n_in + k] : 0.0;
    }
  }
  __syncthreads();
  if (!active) return;
  v4d acc = v4d{0, 0, 0, 0};
  #pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const double b = sB[(ks*4 + (lane >> 4))*BS + (lane & 15)];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[ks], b, acc, 0, 0, 0);
  }
  // register i of lane l: row (l/16) + 4i of column l%16
  const int col = lane & 15;
  if (partial) {
    double *dst = partial + (((size_t)g*8 + o)*GW + col)*a.n_out;
    for (int i = 0; i < 4; ++i) {
      const int row = rb*16 + (lane >> 4) + 4*i;
      if (row < a.n_out) dst[row] = acc[i];
    }
  } else {
    const int tb = a.grp_targets[(size_t)g*GW + col];
    if (tb >= 0 && inputs[col] >= 0) {
      double *dst = a.tgt + (size_t)(tb - a.tbase)*a.n_out;
      for (int i = 0; i < 4; ++i) {
        const int row = rb*16 + (lane >> 4) + 4*i;
        if (row < a.n_out) dst[row] += acc[i];
      }
    }
  }
}

// tgt[target] += sum over the group's octants (in order) of the partial sums
__global__ __launch_bounds__(256) void octant_reduce(OctArgs a, const double *partial) {
  const size_t n = (size_t)a.ngroups*GW*a.n_out;
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x*blockDim.x) {
    const size_t gj = i/a.n_out;
    const int row = i - gj*a.n_out;
    const int g = gj/GW, j = gj - (size_t)g*GW;
    const int tb = a.grp_targets[gj];
    if (tb < 0) continue;
    const int mask = a.grp_mask[g];
    double s = 0;
    for (int o = 0; o < 8; ++o)
      if ((mask >> o) & 1) s += partial[(((size_t)g*8 + o)*GW + j)*a.n_out + row];
    a.tgt[(size_t)(tb - a.tbase)*a.n_out + row] += s;
  }
}

template <int RBW>
int launch_octant(const OctArgs &a, hipStream_t st) {
  const size_t lds = sizeof(double)*(size_t)a.ksteps*4*BS;
  if (lds > 64*1024) {
    hipError_t e = hipFuncSetAttribute((const void *)octant_translate_kernel<RBW>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
      return set_error(SUMPYHIP_EHIP, "octant_translate: cannot raise LDS limit: %s",
                       hipGetErrorString(e));
  }
  const int grid = a.ngroups < (1 << 30) ? a.ngroups : (1 << 30);
  octant_translate_kernel<RBW><<<grid, 64*NWAVE, lds, st>>>(a);
  return check_launch("octant_translate");
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" int sumpyhip_octant_translate(int ngroups, const int *group_targets,
                                         const int *group_inputs, const int *group_octant_mask,
                                         int nitems, const int *items, int n_out, int n_in,
                                         const double *T_frag,
                                         const double *src_expansions, int src_base_ibox,
                                         double *tgt_expansions, int tgt_base_ibox,
                                         void *partial, void *stream) {
  SUMPY_REQUIRE(ngroups >= 0 && n_out > 0 && n_in > 0, "octant_translate: bad sizes");
  SUMPY_REQUIRE(n_in <= 1024, "octant_translate: n_in %d > 1024", n_in);
  if (ngroups == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(group_targets && group_inputs && group_octant_mask && T_frag &&
                    src_expansions && tgt_expansions,
                "octant_translate: null pointer");
  OctArgs a{ngroups, group_targets, group_inputs, group_octant_mask, n_out, n_in,
            (n_out + 15)/16, (n_in + 3)/4, T_frag, src_expansions, src_base_ibox,
            tgt_expansions, tgt_base_ibox};
  hipStream_t st = (hipStream_t)stream;
  const int rbw = (a.nrb + NWAVE - 1)/NWAVE;
  if (rbw == 1 && a.ksteps <= KS) {
    SUMPY_REQUIRE(items && nitems >= 0 && nitems <= 8*ngroups,
                  "octant_translate: %d items for %d groups", nitems, ngroups);
    if (nitems == 0) return SUMPYHIP_OK;
    const size_t lds = sizeof(double)*(size_t)KS*4*BS;
    double *part = static_cast<double *>(partial);
    octant_translate_split<<<nitems, 64*NWAVE, lds, st>>>(a, items, part);
    int rc = check_launch("octant_translate");
    if (rc || !part) return rc;
    const size_t n = (size_t)ngroups*GW*n_out;
    const int grid = (int)((n + 255)/256 < 65536 ? (n + 255)/256 : 65536);
    octant_reduce<<<grid, 256, 0, st>>>(a, part);
    return check_launch("octant_reduce");
  }
  switch (rbw) {
    case 1: return launch_octant<1>(a, st);
    case 2: return launch_octant<2>(a, st);
    case 3: return launch_octant<3>(a, st);
    case 4: return launch_octant<4>(a, st);
    default:
      return set_error(SUMPYHIP_ENOTIMPL, "octant_translate: n_out %d > 512", n_out);
  }
}
