// Device-side octree / quadtree and FMM traversal builder.
//
// The reference takes its tree, interaction lists and translation classes
// from boxtree (TreeBuilder, FMMTraversalBuilder, TranslationClassesBuilder;
// called at sumpy/test/test_fmm.py:182-196 and read at sumpy/fmm.py:
// 357-592).  This file builds the same data model as the host builder
// (tree.cpp) -- the same box numbering, particle order and lists, array for
// array -- with the work on the GPU:
//   * Morton keys of the particles (then a stable sort by the caller);
//   * the boxes level by level: every box split holds more than
//     max_particles_in_box particles (sources + targets); a split box's
//     children are the non-empty cells of its particle ranges, found by
//     binary search in the sorted keys, numbered level by level in Morton
//     order (parents in order, children in child order);
//   * colleagues by binary search of the neighbouring cells' Morton prefixes
//     in the level's (sorted) box prefixes;
//   * lists 1-4 per box by the host builder's depth-first walks (one thread
//     per box, an explicit stack), counted in a first pass and written at
//     scanned offsets in a second, each box's list sorted like the host's;
//   * translation-class slots of the list-2 pairs.
// The orchestration (allocation, scans, compaction) is in sumpy_amd/tree.py.
#include "common.h"

#include <cstdint>

namespace sumpyhip {
namespace {

constexpr int TB = 256;
constexpr int STACK = 192;  // DFS stack: <= 7 pending siblings per level, 21+ levels

__device__ __forceinline__ uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

__device__ __forceinline__ uint64_t spread2(uint64_t v) {
  v &= 0xffffffffULL;
  v = (v | v << 16) & 0x0000ffff0000ffffULL;
  v = (v | v << 8) & 0x00ff00ff00ff00ffULL;
  v = (v | v << 4) & 0x0f0f0f0f0f0f0f0fULL;
  v = (v | v << 2) & 0x3333333333333333ULL;
  v = (v | v << 1) & 0x5555555555555555ULL;
  return v;
}

// Morton key of a cell given by integer coordinates (bit b of axis d goes to
// bit dim b + dim - 1 - d)
__device__ __forceinline__ uint64_t cell_key(int dim, const int64_t *ix) {
  uint64_t k = 0;
  for (int d = 0; d < dim; ++d)
    k |= (dim == 3 ? spread3((uint64_t)ix[d]) : spread2((uint64_t)ix[d])) << (dim - 1 - d);
  return k;
}

// the particles' keys: u = (x - lo) / ext, c = floor(u 2^maxdepth) clamped
__global__ void morton_kernel(int dim, long n, const double *x, const double *y,
                              const double *z, double lo0, double lo1, double lo2, double ext,
                              int maxdepth, uint64_t *keys) {
#pragma clang fp contract(off)
  const uint64_t nbits = (uint64_t)1 << maxdepth;
  const double *X[3] = {x, y, z};
  const double lo[3] = {lo0, lo1, lo2};
  for (long i = blockIdx.x*(long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x*blockDim.x) {
    uint64_t k = 0;
    for (int d = 0; d < dim; ++d) {
      const double u = (X[d][i] - lo[d])/ext;
      int64_t c = (int64_t)floor(u*(double)nbits);
      c = c < 0 ? 0 : (c > (int64_t)nbits - 1 ? (int64_t)nbits - 1 : c);
      k |= (dim == 3 ? spread3((uint64_t)c) : spread2((uint64_t)c)) << (dim - 1 - d);
    }
    keys[i] = k;
  }
}

// first index in [lo, hi) whose key prefix (key >> shift) is >= p
__device__ __forceinline__ long lower_bound_prefix(const uint64_t *keys, long lo, long hi,
                                                   uint64_t p, int shift) {
  while (lo < hi) {
    const long mid = lo + ((hi - lo) >> 1);
    if ((keys[mid] >> shift) < p) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Children of the split boxes of one level: for split box j (prefix P, source
// range [slo, shi), target range [tlo, thi)) and child c, the ranges of the
// prefix (P << dim) | c at the child level (shift = dim (maxdepth - level - 1)).
__global__ void children_kernel(int dim, int nsplit, const uint64_t *prefix, const int *slo,
                                const int *shi, const int *tlo, const int *thi,
                                const uint64_t *skeys, const uint64_t *tkeys, int shift,
                                int *cslo, int *cshi, int *ctlo, int *cthi, int *exists) {
  const int nch = 1 << dim;
  for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < (long)nsplit*nch;
       e += (long)gridDim.x*blockDim.x) {
    const int j = (int)(e/nch), c = (int)(e - (long)j*nch);
    const uint64_t p = (prefix[j] << dim) | (uint64_t)c;
    const long s0 = lower_bound_prefix(skeys, slo[j], shi[j], p, shift);
    const long s1 = lower_bound_prefix(skeys, s0, shi[j], p + 1, shift);
    long t0 = s0, t1 = s1;
    if (tkeys) {
      t0 = lower_bound_prefix(tkeys, tlo[j], thi[j], p, shift);
      t1 = lower_bound_prefix(tkeys, t0, thi[j], p + 1, shift);
    }
    cslo[e] = (int)s0; cshi[e] = (int)s1; ctlo[e] = (int)t0; cthi[e] = (int)t1;
    exists[e] = (s1 > s0 || t1 > t0) ? 1 : 0;
  }
}

// integer cell coordinates of a box from its Morton prefix at its level
__device__ __forceinline__ void prefix_to_ix(int dim, uint64_t p, int level, int64_t *ix) {
  for (int d = 0; d < dim; ++d) ix[d] = 0;
  for (int b = 0; b < level; ++b)
    for (int d = 0; d < dim; ++d)
      ix[d] |= (int64_t)((p >> (dim*b + dim - 1 - d)) & 1) << b;
}

// box centres (lo + (ix + 1/2) h, h = ext / 2^level, multiply then add as two
// separately rounded operations, as the host builder computes them) and
// integer coordinates
__global__ void box_geometry_kernel(int dim, int nboxes, const uint64_t *prefix,
                                    const int *level, double lo0, double lo1, double lo2,
                                    double ext, double *centers, int *ixo) {
#pragma clang fp contract(off)
  const double lo[3] = {lo0, lo1, lo2};
  for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < nboxes; i += gridDim.x*blockDim.x) {
    int64_t ix[3];
    prefix_to_ix(dim, prefix[i], level[i], ix);
    const double h = ext/(double)((int64_t)1 << level[i]);
    for (int d = 0; d < dim; ++d) {
      const double off = ((double)ix[d] + 0.5)*h;
      centers[(size_t)d*nboxes + i] = lo[d] + off;
      ixo[(size_t)d*nboxes + i] = (int)ix[d];
    }
  }
}

struct TreeView {
  int dim, nboxes, nlevels;
  const int *level, *ix, *child, *parent, *leaf, *level_start;
  const int *src_cumul, *src_nonchild;
  const uint64_t *prefix;
};

__device__ __forceinline__ bool adjacent(const TreeView &t, int a, int b) {
  const int la = t.level[a], lb = t.level[b];
  const int L = la > lb ? la : lb;
  for (int d = 0; d < t.dim; ++d) {
    const int64_t ax = t.ix[(size_t)d*t.nboxes + a], bx = t.ix[(size_t)d*t.nboxes + b];
    const int64_t alo = ax << (L - la), ahi = (ax + 1) << (L - la);
    const int64_t blo = bx << (L - lb), bhi = (bx + 1) << (L - lb);
    if (alo > bhi || blo > ahi) return false;
  }
  return true;
}

__device__ __forceinline__ int child_of(const TreeView &t, int b, int c) {
  return t.child[(size_t)c*t.nboxes + b];
}

// the box at level l with cell coordinates ix, or -1
__device__ int find_box(const TreeView &t, int l, const int64_t *ix) {
  for (int d = 0; d < t.dim; ++d)
    if (ix[d] < 0 || ix[d] >= ((int64_t)1 << l)) return -1;
  const uint64_t key = cell_key(t.dim, ix);
  int lo = t.level_start[l], hi = t.level_start[l + 1];
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (t.prefix[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < t.level_start[l + 1] && t.prefix[lo] == key) ? lo : -1;
}

// colleagues (same-level neighbours) of every box, ascending, -1 padded
__global__ void colleagues_kernel(TreeView t, int *coll) {
  const int ncol = t.dim == 3 ? 26 : 8;
  for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < t.nboxes; i += gridDim.x*blockDim.x) {
    int found[26];
    int n = 0;
    const int l = t.level[i];
    const int ncomb = t.dim == 3 ? 27 : 9;
    for (int c = 0; c < ncomb; ++c) {
      int cc = c;
      bool zero = true;
      int64_t ix[3];
      for (int d = 0; d < t.dim; ++d) {
        const int off = cc % 3 - 1;
        cc /= 3;
        if (off) zero = false;
        ix[d] = t.ix[(size_t)d*t.nboxes + i] + off;
      }
      if (zero) continue;
      const int j = find_box(t, l, ix);
      if (j >= 0) {
        int k = n++;  // insertion, ascending
        while (k > 0 && found[k - 1] > j) { found[k] = found[k - 1]; --k; }
        found[k] = j;
      }
    }
    for (int k = 0; k < ncol; ++k) coll[(size_t)i*ncol + k] = k < n ? found[k] : -1;
  }
}

__device__ __forceinline__ void sort_segment(int *a, int n) {
  for (int i = 1; i < n; ++i) {
    const int v = a[i];
    int k = i;
    while (k > 0 && a[k - 1] > v) { a[k] = a[k - 1]; --k; }
    a[k] = v;
  }
}

// List 1: leaf source boxes adjacent to each target box (host: depth-first
// from the root, children pushed in reverse order).  counts != null: the
// counting pass; else the writing pass (out at starts).  The same convention
// holds for every list kernel below.
__global__ void list1_kernel(TreeView t, int nt, const int *target_boxes, int *counts,
                             const int *starts, int *out) {
  const int nch = 1 << t.dim;
  for (int it = blockIdx.x*blockDim.x + threadIdx.x; it < nt; it += gridDim.x*blockDim.x) {
    const int tb = target_boxes[it];
    int stack[STACK];
    int sp = 0, n = 0;
    const bool counting = counts != nullptr;
    int *o = counting ? nullptr : out + starts[it];
    stack[sp++] = 0;
    while (sp) {
      const int x = stack[--sp];
      if (!adjacent(t, x, tb)) continue;
      if (t.leaf[x]) {
        if (t.src_nonchild[x] > 0) { if (!counting) o[n] = x; ++n; }
        continue;
      }
      if (t.src_cumul[x] <= 0) continue;
      for (int c = nch - 1; c >= 0; --c) {
        const int ch = child_of(t, x, c);
        if (ch && sp < STACK) stack[sp++] = ch;
      }
    }
    if (counting) counts[it] = n; else sort_segment(o, n);
  }
}

// List 2: children with sources of the parent's colleagues that are not
// adjacent to the box (ascending: colleagues ascending, children in order);
// with the translation-class slot of each pair and the level's used slots.
__global__ void list2_kernel(TreeView t, int nb, const int *boxes, const int *coll,
                             int *counts, const int *starts, int *out, int *slot_out,
                             int *used) {
  const int nch = 1 << t.dim, ncol = t.dim == 3 ? 26 : 8, nslots = t.dim == 3 ? 343 : 49;
  for (int ib = blockIdx.x*blockDim.x + threadIdx.x; ib < nb; ib += gridDim.x*blockDim.x) {
    const int b = boxes[ib];
    const int l = t.level[b];
    int n = 0;
    if (l >= 2) {
      const int p = t.parent[b];
      for (int k = 0; k < ncol; ++k) {
        const int c = coll[(size_t)p*ncol + k];
        if (c < 0) break;
        for (int q = 0; q < nch; ++q) {
          const int ch = child_of(t, c, q);
          if (!ch || t.src_cumul[ch] <= 0) continue;
          if (adjacent(t, ch, b)) continue;
          if (!counts) {
            const int e = starts[ib] + n;
            out[e] = ch;
            int slot = 0;
            for (int d = 0; d < t.dim; ++d)
              slot = slot*7 + (int)(t.ix[(size_t)d*t.nboxes + b] - t.ix[(size_t)d*t.nboxes + ch] + 3);
            slot_out[e] = slot;
            // most pairs find their slot already flagged: read before the
            // atomic (a stale 0 only costs one more atomic)
            int *u = &used[(size_t)l*nslots + slot];
            if (*u == 0) atomicOr(u, 1);
          }
          ++n;
        }
      }
    }
    if (counts) counts[ib] = n;
  }
}

// List 3: descendants with sources of the colleagues that are separated from
// the target box while their parents are adjacent to it, grouped by source
// level.  Count pass: counts[it * nlevels + level]; write pass: entry of
// (level, target) at offsets[level * nt + it], each group sorted.
__global__ void list3_kernel(TreeView t, int nt, const int *target_boxes, const int *coll,
                             int *counts, const int *offsets, int *out) {
  const int nch = 1 << t.dim, ncol = t.dim == 3 ? 26 : 8;
  for (int it = blockIdx.x*blockDim.x + threadIdx.x; it < nt; it += gridDim.x*blockDim.x) {
    const int tb = target_boxes[it];
    int fill[32];
    for (int l = 0; l < t.nlevels && l < 32; ++l) fill[l] = 0;
    int stack[STACK];
    for (int k = 0; k < ncol; ++k) {
      const int c = coll[(size_t)tb*ncol + k];
      if (c < 0) break;
      if (t.leaf[c]) continue;
      int sp = 0;
      stack[sp++] = c;
      while (sp) {
        const int x = stack[--sp];
        for (int q = 0; q < nch; ++q) {
          const int ch = child_of(t, x, q);
          if (!ch || t.src_cumul[ch] <= 0) continue;
          if (adjacent(t, ch, tb)) {
            if (!t.leaf[ch] && sp < STACK) stack[sp++] = ch;
          } else {
            const int ls = t.level[ch];
            if (!counts) out[offsets[(size_t)ls*nt + it] + fill[ls]] = ch;
            ++fill[ls];
          }
        }
      }
    }
    for (int l = 0; l < t.nlevels && l < 32; ++l) {
      if (counts) counts[(size_t)it*t.nlevels + l] = fill[l];
      else if (fill[l]) sort_segment(out + offsets[(size_t)l*nt + it], fill[l]);
    }
  }
}

// List 4: leaves with sources, adjacent to the parent (not the parent itself,
// no deeper than the parent's level), separated from the box; ascending.
__global__ void list4_kernel(TreeView t, int nb, const int *boxes, int *counts,
                             const int *starts, int *out) {
  const int nch = 1 << t.dim;
  for (int ib = blockIdx.x*blockDim.x + threadIdx.x; ib < nb; ib += gridDim.x*blockDim.x) {
    const int b = boxes[ib];
    int n = 0;
    const bool counting = counts != nullptr;
    int *o = counting ? nullptr : out + starts[ib];
    if (t.level[b] >= 1) {
      const int p = t.parent[b], lp = t.level[p];
      int stack[STACK];
      int sp = 0;
      stack[sp++] = 0;
      while (sp) {
        const int x = stack[--sp];
        if (x == p || !adjacent(t, x, p)) continue;
        if (t.leaf[x]) {
          if (t.src_nonchild[x] > 0 && !adjacent(t, x, b)) { if (!counting) o[n] = x; ++n; }
          continue;
        }
        if (t.level[x] >= lp || t.src_cumul[x] <= 0) continue;
        for (int c = 0; c < nch; ++c) {
          const int ch = child_of(t, x, c);
          if (ch && sp < STACK) stack[sp++] = ch;
        }
      }
    }
    if (counting) counts[ib] = n; else sort_segment(o, n);
  }
}

// translation class of each list-2 entry: cls[e] = slot_to_cls[level(box)][slot[e]]
__global__ void list2_classes_kernel(int n, const int *entry_level, const int *slot,
                                     const int *slot_to_cls, int nslots, int *cls) {
  for (int e = blockIdx.x*blockDim.x + threadIdx.x; e < n; e += gridDim.x*blockDim.x)
    cls[e] = slot_to_cls[(size_t)entry_level[e]*nslots + slot[e]];
}

int grid_of(long n) {
  const long g = (n + TB - 1)/TB;
  return (int)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

TreeView make_view(int dim, int nboxes, int nlevels, const int *level, const int *ix,
                   const int *child, const int *parent, const int *leaf,
                   const int *level_start, const int *src_cumul, const int *src_nonchild,
                   const uint64_t *prefix) {
  TreeView t;
  t.dim = dim; t.nboxes = nboxes; t.nlevels = nlevels;
  t.level = level; t.ix = ix; t.child = child; t.parent = parent; t.leaf = leaf;
  t.level_start = level_start; t.src_cumul = src_cumul; t.src_nonchild = src_nonchild;
  t.prefix = prefix;
  return t;
}

}  // namespace
}  // namespace sumpyhip

using namespace sumpyhip;

extern "C" int sumpytree_gpu_morton(int dim, long n, const double *x, const double *y,
                                    const double *z, const double *lo, double ext,
                                    int maxdepth, void *keys, void *stream) {
  SUMPY_REQUIRE(dim == 2 || dim == 3, "tree_gpu: dim must be 2 or 3");
  SUMPY_REQUIRE(maxdepth >= 1 && maxdepth <= (dim == 3 ? 21 : 31), "tree_gpu: bad depth");
  if (n == 0) return SUMPYHIP_OK;
  morton_kernel<<<grid_of(n), TB, 0, (hipStream_t)stream>>>(
      dim, n, x, y, dim == 3 ? z : x, lo[0], lo[1], dim == 3 ? lo[2] : 0.0, ext, maxdepth,
      (uint64_t *)keys);
  return check_launch("tree_gpu_morton");
}

extern "C" int sumpytree_gpu_children(int dim, int nsplit, const void *prefix, const int *slo,
                                      const int *shi, const int *tlo, const int *thi,
                                      const void *skeys, const void *tkeys, int shift,
                                      int *cslo, int *cshi, int *ctlo, int *cthi, int *exists,
                                      void *stream) {
  if (nsplit == 0) return SUMPYHIP_OK;
  children_kernel<<<grid_of((long)nsplit << dim), TB, 0, (hipStream_t)stream>>>(
      dim, nsplit, (const uint64_t *)prefix, slo, shi, tlo, thi, (const uint64_t *)skeys,
      (const uint64_t *)tkeys, shift, cslo, cshi, ctlo, cthi, exists);
  return check_launch("tree_gpu_children");
}

extern "C" int sumpytree_gpu_geometry(int dim, int nboxes, const void *prefix,
                                      const int *level, const double *lo, double ext,
                                      double *centers, int *ix, void *stream) {
  if (nboxes == 0) return SUMPYHIP_OK;
  box_geometry_kernel<<<grid_of(nboxes), TB, 0, (hipStream_t)stream>>>(
      dim, nboxes, (const uint64_t *)prefix, level, lo[0], lo[1], dim == 3 ? lo[2] : 0.0, ext,
      centers, ix);
  return check_launch("tree_gpu_geometry");
}

extern "C" int sumpytree_gpu_colleagues(int dim, int nboxes, int nlevels, const int *level,
                                        const int *ix, const int *child, const int *parent,
                                        const int *leaf, const int *level_start,
                                        const int *src_cumul, const int *src_nonchild,
                                        const void *prefix, int *coll, void *stream) {
  SUMPY_REQUIRE(nlevels <= 32, "tree_gpu: at most 32 levels");
  TreeView t = make_view(dim, nboxes, nlevels, level, ix, child, parent, leaf, level_start,
                         src_cumul, src_nonchild, (const uint64_t *)prefix);
  colleagues_kernel<<<grid_of(nboxes), TB, 0, (hipStream_t)stream>>>(t, coll);
  return check_launch("tree_gpu_colleagues");
}

// which: 1 (list 1, over target boxes), 2 (list 2, over target-or-parent
// boxes; slots and used), 3 (list 3, over target boxes; counts per level,
// offsets per (level, box)), 4 (list 4, over target-or-parent boxes).
// counts != null: the counting pass (starts, out, slots, used unused); else
// the writing pass (starts, out and, for list 2, slots and used required).
extern "C" int sumpytree_gpu_list(int which, int dim, int nboxes, int nlevels,
                                  const int *level, const int *ix, const int *child,
                                  const int *parent, const int *leaf, const int *level_start,
                                  const int *src_cumul, const int *src_nonchild,
                                  const void *prefix, const int *coll, int nb,
                                  const int *boxes, int *counts, const int *starts, int *out,
                                  int *slots, int *used, void *stream) {
  SUMPY_REQUIRE(which >= 1 && which <= 4, "tree_gpu: list %d", which);
  SUMPY_REQUIRE(nlevels <= 32, "tree_gpu: at most 32 levels");
  if (nb == 0) return SUMPYHIP_OK;
  SUMPY_REQUIRE(counts || (starts && out && (which != 2 || (slots && used))),
                "tree_gpu: list %d writing pass without its output arrays", which);
  TreeView t = make_view(dim, nboxes, nlevels, level, ix, child, parent, leaf, level_start,
                         src_cumul, src_nonchild, (const uint64_t *)prefix);
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_of(nb);
  // the walks keep a private stack: a few boxes per wave in flight at a time
  if (which == 1) list1_kernel<<<g, TB, 0, st>>>(t, nb, boxes, counts, starts, out);
  else if (which == 2)
    list2_kernel<<<g, TB, 0, st>>>(t, nb, boxes, coll, counts, starts, out, slots, used);
  else if (which == 3) list3_kernel<<<g, TB, 0, st>>>(t, nb, boxes, coll, counts, starts, out);
  else list4_kernel<<<g, TB, 0, st>>>(t, nb, boxes, counts, starts, out);
  return check_launch("tree_gpu_list");
}

extern "C" int sumpytree_gpu_list2_classes(int n, const int *entry_level, const int *slot,
                                           const int *slot_to_cls, int nslots, int *cls,
                                           void *stream) {
  if (n == 0) return SUMPYHIP_OK;
  list2_classes_kernel<<<grid_of(n), TB, 0, (hipStream_t)stream>>>(n, entry_level, slot,
                                                                   slot_to_cls, nslots, cls);
  return check_launch("tree_gpu_list2_classes");
}
