// Host-side octree/quadtree and FMM traversal builder.
//
// The reference takes its tree, its interaction lists and its M2L translation
// classes from boxtree (TreeBuilder, FMMTraversalBuilder,
// TranslationClassesBuilder; called at sumpy/test/test_fmm.py:190-196 and
// sumpy/fmm.py:357-363), which is not part of this build.  This file builds
// the same data model with the field names the wrangler reads
// (sumpy/fmm.py:372-592,752-796):
//   * root box: particle bounding box, made cubic with extent
//     max_extent * (1 + 1e-4) anchored at the bounding-box minimum;
//   * a box is split while it holds more than max_particles_in_box particles
//     (sources + targets, or sources when the targets are the sources); empty
//     children are not created;
//   * boxes are numbered level by level, Morton order within a level; the
//     particles are sorted into tree order so that every box's particles are
//     contiguous;
//   * lists 1-4 (U, V, W, X lists of the adaptive FMM) with "adjacent" meaning
//     that the closed boxes touch;
//   * translation classes: per level, the integer offsets (c_tgt - c_src)/h
//     of list-2 pairs, numbered in increasing lexicographic order of the
//     offset among those that occur.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct Box {
  int level;
  int64_t ix[3];
  int parent;
  int child[8];
  bool leaf;
  long src_lo, src_hi, tgt_lo, tgt_hi;
};

struct Result {
  int dim = 3;
  double root_extent = 0;
  double bbox_min[3] = {0, 0, 0};
  int nlevels = 0;
  std::map<std::string, std::vector<int32_t>> ints;
  std::map<std::string, std::vector<double>> dbls;
  std::string error;
};

inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}
inline uint64_t spread2(uint64_t v) {
  v &= 0xffffffffULL;
  v = (v | v << 16) & 0x0000ffff0000ffffULL;
  v = (v | v << 8) & 0x00ff00ff00ff00ffULL;
  v = (v | v << 4) & 0x0f0f0f0f0f0f0f0fULL;
  v = (v | v << 2) & 0x3333333333333333ULL;
  v = (v | v << 1) & 0x5555555555555555ULL;
  return v;
}

struct Builder {
  int dim;
  int maxdepth;
  std::vector<Box> boxes;
  std::vector<std::unordered_map<uint64_t, int>> level_maps;

  uint64_t cell_key(const int64_t *ix) const {
    uint64_t k = 0;
    for (int d = 0; d < dim; ++d) k = (k << 21) | (uint64_t)ix[d];
    return k;
  }
  int find(int level, const int64_t *ix) const {
    if (level >= (int)level_maps.size()) return -1;
    for (int d = 0; d < dim; ++d)
      if (ix[d] < 0 || ix[d] >= ((int64_t)1 << level)) return -1;
    auto it = level_maps[level].find(cell_key(ix));
    return it == level_maps[level].end() ? -1 : it->second;
  }
  // closed boxes touch or overlap
  bool adjacent(int a, int b) const {
    const Box &A = boxes[a], &B = boxes[b];
    int L = std::max(A.level, B.level);
    for (int d = 0; d < dim; ++d) {
      int64_t alo = A.ix[d] << (L - A.level), ahi = (A.ix[d] + 1) << (L - A.level);
      int64_t blo = B.ix[d] << (L - B.level), bhi = (B.ix[d] + 1) << (L - B.level);
      if (alo > bhi || blo > ahi) return false;
    }
    return true;
  }
};

}  // namespace

extern "C" {

void *sumpytree_build_in_box(int dim, long nsrc, const double *src, long ntgt,
                             const double *tgt, int max_particles_in_box, int max_levels,
                             const double *box_min, double box_extent);

// Builds the tree and traversal.  src/tgt are SoA (dim x n, row-major);
// tgt == nullptr means the targets are the sources.
void *sumpytree_build(int dim, long nsrc, const double *src, long ntgt,
                      const double *tgt, int max_particles_in_box, int max_levels) {
  return sumpytree_build_in_box(dim, nsrc, src, ntgt, tgt, max_particles_in_box, max_levels,
                                nullptr, 0);
}

// As sumpytree_build, with the root box given (box_min, cubic extent; every
// particle must lie inside) instead of the particles' bounding box: trees of
// several problems built in one root box share box sizes and centre lattice
// level by level, so they can be merged into one forest.
void *sumpytree_build_in_box(int dim, long nsrc, const double *src, long ntgt,
                             const double *tgt, int max_particles_in_box, int max_levels,
                             const double *box_min, double box_extent) {
  Result *res = new Result;
  res->dim = dim;
  if (dim != 2 && dim != 3) { res->error = "dim must be 2 or 3"; return res; }
  const bool same = (tgt == nullptr);
  if (same) ntgt = nsrc;
  const double *T = same ? src : tgt;
  if (nsrc + (same ? 0 : ntgt) == 0) { res->error = "no particles"; return res; }
  if (max_particles_in_box < 1) { res->error = "max_particles_in_box must be >= 1"; return res; }

  // {{{ bounding box
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int d = 0; d < dim; ++d) {
    for (long i = 0; i < nsrc; ++i) { lo[d] = std::min(lo[d], src[d*nsrc + i]); hi[d] = std::max(hi[d], src[d*nsrc + i]); }
    if (!same) for (long i = 0; i < ntgt; ++i) { lo[d] = std::min(lo[d], T[d*ntgt + i]); hi[d] = std::max(hi[d], T[d*ntgt + i]); }
  }
  double ext = 0;
  for (int d = 0; d < dim; ++d) ext = std::max(ext, hi[d] - lo[d]);
  if (ext == 0) ext = 1;
  ext *= 1 + 1e-4;
  if (box_min) {
    for (int d = 0; d < dim; ++d) {
      if (lo[d] < box_min[d] || hi[d] >= box_min[d] + box_extent) {
        res->error = "particles outside the given root box";
        return res;
      }
      lo[d] = box_min[d];
    }
    ext = box_extent;
  }
  res->root_extent = ext;
  for (int d = 0; d < dim; ++d) res->bbox_min[d] = lo[d];
  // }}}

  const int maxdepth = std::min(max_levels > 0 ? max_levels : 21, dim == 3 ? 21 : 31);
  const uint64_t nbits = (uint64_t)1 << maxdepth;
  auto keys_of = [&](const double *X, long n, std::vector<uint64_t> &keys) {
    keys.resize(n);
    for (long i = 0; i < n; ++i) {
      uint64_t k = 0;
      for (int d = 0; d < dim; ++d) {
        double u = (X[d*n + i] - lo[d])/ext;
        int64_t c = (int64_t)std::floor(u*(double)nbits);
        c = std::max<int64_t>(0, std::min<int64_t>(c, nbits - 1));
        k |= (dim == 3 ? spread3(c) : spread2(c)) << (dim - 1 - d);
      }
      keys[i] = k;
    }
  };
  std::vector<uint64_t> skeys, tkeys;
  keys_of(src, nsrc, skeys);
  if (!same) keys_of(T, ntgt, tkeys);
  std::vector<int32_t> sperm(nsrc), tperm;
  for (long i = 0; i < nsrc; ++i) sperm[i] = i;
  std::stable_sort(sperm.begin(), sperm.end(), [&](int a, int b) { return skeys[a] < skeys[b]; });
  std::vector<uint64_t> sk_sorted(nsrc), tk_sorted;
  for (long i = 0; i < nsrc; ++i) sk_sorted[i] = skeys[sperm[i]];
  if (!same) {
    tperm.resize(ntgt);
    for (long i = 0; i < ntgt; ++i) tperm[i] = i;
    std::stable_sort(tperm.begin(), tperm.end(), [&](int a, int b) { return tkeys[a] < tkeys[b]; });
    tk_sorted.resize(ntgt);
    for (long i = 0; i < ntgt; ++i) tk_sorted[i] = tkeys[tperm[i]];
  }

  // {{{ top-down build (breadth first)
  Builder B;
  B.dim = dim;
  B.maxdepth = maxdepth;
  const int nch = 1 << dim;
  Box root{};
  root.level = 0; root.parent = 0; root.leaf = true;
  for (int c = 0; c < 8; ++c) root.child[c] = 0;
  root.src_lo = 0; root.src_hi = nsrc;
  root.tgt_lo = 0; root.tgt_hi = same ? nsrc : ntgt;
  B.boxes.push_back(root);
  std::vector<int> frontier = {0};
  int level = 0;
  while (!frontier.empty()) {
    std::vector<int> next;
    for (int b : frontier) {
      Box bx = B.boxes[b];
      long cnt = (bx.src_hi - bx.src_lo) + (same ? 0 : (bx.tgt_hi - bx.tgt_lo));
      if (cnt <= max_particles_in_box || level + 1 >= maxdepth) continue;
      const int shift = dim*(maxdepth - level - 1);
      long s = bx.src_lo, t = bx.tgt_lo;
      for (int c = 0; c < nch; ++c) {
        long s2 = s;
        while (s2 < bx.src_hi && (int)((sk_sorted[s2] >> shift) & (nch - 1)) == c) ++s2;
        long t2 = t;
        if (!same) while (t2 < bx.tgt_hi && (int)((tk_sorted[t2] >> shift) & (nch - 1)) == c) ++t2;
        else t2 = s2;
        if (s2 > s || t2 > t) {
          Box ch{};
          ch.level = level + 1;
          ch.parent = b;
          ch.leaf = true;
          for (int cc = 0; cc < 8; ++cc) ch.child[cc] = 0;
          for (int d = 0; d < dim; ++d)
            ch.ix[d] = 2*bx.ix[d] + ((c >> (dim - 1 - d)) & 1);
          ch.src_lo = s; ch.src_hi = s2;
          ch.tgt_lo = same ? s : t; ch.tgt_hi = same ? s2 : t2;
          int id = (int)B.boxes.size();
          B.boxes.push_back(ch);
          B.boxes[b].child[c] = id;
          B.boxes[b].leaf = false;
          next.push_back(id);
        }
        s = s2; t = same ? s2 : t2;
      }
    }
    frontier.swap(next);
    ++level;
  }
  // }}}

  // boxes are created level by level, each level in Morton order already
  const int nboxes = (int)B.boxes.size();
  int nlevels = 0;
  for (auto &bx : B.boxes) nlevels = std::max(nlevels, bx.level + 1);
  res->nlevels = nlevels;
  B.level_maps.resize(nlevels);
  for (int i = 0; i < nboxes; ++i) B.level_maps[B.boxes[i].level][B.cell_key(B.boxes[i].ix)] = i;

  // {{{ tree arrays
  auto &I = res->ints;
  auto &D = res->dbls;
  std::vector<int32_t> level_start(nlevels + 1, 0);
  for (auto &bx : B.boxes) level_start[bx.level + 1]++;
  for (int l = 0; l < nlevels; ++l) level_start[l + 1] += level_start[l];
  I["level_start_box_nrs"] = level_start;
  std::vector<int32_t> bss(nboxes), bscc(nboxes), bscn(nboxes), bts(nboxes), btcc(nboxes),
      btcn(nboxes), parent(nboxes), levels(nboxes), child(nch*(size_t)nboxes, 0), flags(nboxes);
  std::vector<double> centers(dim*(size_t)nboxes);
  for (int i = 0; i < nboxes; ++i) {
    const Box &bx = B.boxes[i];
    bss[i] = bx.src_lo; bscc[i] = bx.src_hi - bx.src_lo; bscn[i] = bx.leaf ? bscc[i] : 0;
    bts[i] = bx.tgt_lo; btcc[i] = bx.tgt_hi - bx.tgt_lo; btcn[i] = bx.leaf ? btcc[i] : 0;
    parent[i] = bx.parent; levels[i] = bx.level;
    for (int c = 0; c < nch; ++c) child[(size_t)c*nboxes + i] = bx.child[c];
    double h = ext/(double)((int64_t)1 << bx.level);
    for (int d = 0; d < dim; ++d) centers[(size_t)d*nboxes + i] = lo[d] + (bx.ix[d] + 0.5)*h;
    // boxtree box flags: 1 own srcs, 2 own tgts, 4 child srcs, 8 child tgts
    int f = 0;
    if (bscn[i]) f |= 1;
    if (btcn[i]) f |= 2;
    if (!bx.leaf && bscc[i]) f |= 4;
    if (!bx.leaf && btcc[i]) f |= 8;
    flags[i] = f;
  }
  I["box_source_starts"] = bss; I["box_source_counts_cumul"] = bscc;
  I["box_source_counts_nonchild"] = bscn; I["box_target_starts"] = bts;
  I["box_target_counts_cumul"] = btcc; I["box_target_counts_nonchild"] = btcn;
  I["box_parent_ids"] = parent; I["box_levels"] = levels; I["box_child_ids"] = child;
  I["box_flags"] = flags;
  D["box_centers"] = centers;
  I["user_source_ids"] = sperm;
  {
    std::vector<int32_t> sti(ntgt);
    const std::vector<int32_t> &tp = same ? sperm : tperm;
    for (long i = 0; i < ntgt; ++i) sti[tp[i]] = i;
    I["sorted_target_ids"] = sti;
    if (!same) I["user_target_ids"] = tperm;
  }
  // }}}

  // {{{ box classes
  auto has_own_src = [&](int b) { return bscn[b] > 0; };
  auto has_src = [&](int b) { return bscc[b] > 0; };
  std::vector<int32_t> source_boxes, source_parent_boxes, target_boxes, totp_boxes;
  std::vector<int32_t> ls_src(nlevels + 1, 0), ls_srcpar(nlevels + 1, 0), ls_tgt(nlevels + 1, 0),
      ls_totp(nlevels + 1, 0);
  for (int i = 0; i < nboxes; ++i) {
    int l = B.boxes[i].level;
    if (bscn[i]) { source_boxes.push_back(i); ls_src[l + 1]++; }
    if (!B.boxes[i].leaf && bscc[i]) { source_parent_boxes.push_back(i); ls_srcpar[l + 1]++; }
    if (btcn[i]) { target_boxes.push_back(i); ls_tgt[l + 1]++; }
    if (btcc[i]) { totp_boxes.push_back(i); ls_totp[l + 1]++; }
  }
  for (int l = 0; l < nlevels; ++l) {
    ls_src[l + 1] += ls_src[l]; ls_srcpar[l + 1] += ls_srcpar[l];
    ls_tgt[l + 1] += ls_tgt[l]; ls_totp[l + 1] += ls_totp[l];
  }
  I["source_boxes"] = source_boxes; I["level_start_source_box_nrs"] = ls_src;
  I["source_parent_boxes"] = source_parent_boxes;
  I["level_start_source_parent_box_nrs"] = ls_srcpar;
  I["target_boxes"] = target_boxes; I["level_start_target_box_nrs"] = ls_tgt;
  I["target_or_target_parent_boxes"] = totp_boxes;
  I["level_start_target_or_target_parent_box_nrs"] = ls_totp;
  // }}}

  // {{{ colleagues
  std::vector<std::vector<int>> colleagues(nboxes);
  for (int i = 0; i < nboxes; ++i) {
    const Box &bx = B.boxes[i];
    int64_t off[3];
    int ncomb = dim == 3 ? 27 : 9;
    for (int c = 0; c < ncomb; ++c) {
      int cc = c;
      bool zero = true;
      for (int d = 0; d < dim; ++d) { off[d] = cc % 3 - 1; cc /= 3; if (off[d]) zero = false; }
      if (zero) continue;
      int64_t ix[3];
      for (int d = 0; d < dim; ++d) ix[d] = bx.ix[d] + off[d];
      int j = B.find(bx.level, ix);
      if (j >= 0) colleagues[i].push_back(j);
    }
    std::sort(colleagues[i].begin(), colleagues[i].end());
  }
  // }}}

  // {{{ list 1: leaf source boxes adjacent to each target box
  std::vector<int32_t> l1s(1, 0), l1l;
  {
    std::vector<int> stack;
    for (int t : target_boxes) {
      stack.assign(1, 0);
      std::vector<int> found;
      while (!stack.empty()) {
        int x = stack.back(); stack.pop_back();
        if (!B.adjacent(x, t)) continue;
        if (B.boxes[x].leaf) { if (has_own_src(x)) found.push_back(x); continue; }
        if (!has_src(x)) continue;
        for (int c = nch - 1; c >= 0; --c) if (B.boxes[x].child[c]) stack.push_back(B.boxes[x].child[c]);
      }
      std::sort(found.begin(), found.end());
      l1l.insert(l1l.end(), found.begin(), found.end());
      l1s.push_back((int32_t)l1l.size());
    }
  }
  I["neighbor_source_boxes_starts"] = l1s; I["neighbor_source_boxes_lists"] = l1l;
  // }}}

  // {{{ list 2: well-separated children of the parent's colleagues
  std::vector<int32_t> l2s(1, 0), l2l, l2cls;
  std::vector<std::vector<int>> used_slots(nlevels);
  std::vector<int32_t> l2slot;
  for (int b : totp_boxes) {
    const Box &bx = B.boxes[b];
    std::vector<int> found;
    if (bx.level >= 2) {
      int p = bx.parent;
      for (int c : colleagues[p]) {
        for (int k = 0; k < nch; ++k) {
          int ch = B.boxes[c].child[k];
          if (!ch || !has_src(ch)) continue;
          if (!B.adjacent(ch, b)) found.push_back(ch);
        }
      }
    }
    std::sort(found.begin(), found.end());
    for (int s : found) {
      int slot = 0;
      for (int d = 0; d < dim; ++d) slot = slot*7 + (int)(bx.ix[d] - B.boxes[s].ix[d] + 3);
      l2slot.push_back(slot);
    }
    l2l.insert(l2l.end(), found.begin(), found.end());
    l2s.push_back((int32_t)l2l.size());
  }
  // translation classes, numbered per level
  {
    const int nslots = dim == 3 ? 343 : 49;
    std::vector<std::vector<int>> slot_to_cls(nlevels, std::vector<int>(nslots, -1));
    std::vector<std::vector<char>> used(nlevels, std::vector<char>(nslots, 0));
    size_t e = 0;
    for (size_t ib = 0; ib < totp_boxes.size(); ++ib) {
      int lvl = B.boxes[totp_boxes[ib]].level;
      for (int k = l2s[ib]; k < l2s[ib + 1]; ++k, ++e) used[lvl][l2slot[k]] = 1;
    }
    std::vector<int32_t> cls_ls(nlevels + 1, 0);
    std::vector<double> vecs;
    int gid = 0;
    std::vector<std::vector<int>> vec_of;
    for (int l = 0; l < nlevels; ++l) {
      cls_ls[l] = gid;
      double h = ext/(double)((int64_t)1 << l);
      for (int s = 0; s < nslots; ++s) {
        if (!used[l][s]) continue;
        slot_to_cls[l][s] = gid++;
        int rem = s;
        int v[3];
        for (int d = dim - 1; d >= 0; --d) { v[d] = rem % 7 - 3; rem /= 7; }
        for (int d = 0; d < dim; ++d) vecs.push_back(v[d]*h);
      }
    }
    cls_ls[nlevels] = gid;
    // transpose to (dim, nclasses)
    std::vector<double> vt(dim*(size_t)gid);
    for (int c = 0; c < gid; ++c)
      for (int d = 0; d < dim; ++d) vt[(size_t)d*gid + c] = vecs[(size_t)c*dim + d];
    l2cls.resize(l2l.size());
    for (size_t ib = 0; ib < totp_boxes.size(); ++ib) {
      int lvl = B.boxes[totp_boxes[ib]].level;
      for (int k = l2s[ib]; k < l2s[ib + 1]; ++k) l2cls[k] = slot_to_cls[lvl][l2slot[k]];
    }
    I["from_sep_siblings_translation_classes"] = l2cls;
    I["from_sep_siblings_translation_classes_level_starts"] = cls_ls;
    D["from_sep_siblings_translation_class_to_distance_vector"] = vt;
  }
  I["from_sep_siblings_starts"] = l2s; I["from_sep_siblings_lists"] = l2l;
  // }}}

  // {{{ list 3: separated descendants of colleagues, grouped by source level
  {
    std::vector<std::vector<std::pair<int, int>>> per_level(nlevels);  // (tgt box, src box)
    std::vector<int> stack;
    for (int t : target_boxes) {
      std::vector<int> found;
      for (int c : colleagues[t]) {
        if (B.boxes[c].leaf) continue;
        stack.assign(1, c);
        while (!stack.empty()) {
          int x = stack.back(); stack.pop_back();
          for (int k = 0; k < nch; ++k) {
            int ch = B.boxes[x].child[k];
            if (!ch || !has_src(ch)) continue;
            if (B.adjacent(ch, t)) { if (!B.boxes[ch].leaf) stack.push_back(ch); }
            else found.push_back(ch);
          }
        }
      }
      std::sort(found.begin(), found.end());
      for (int s : found) per_level[B.boxes[s].level].push_back({t, s});
    }
    for (int l = 0; l < nlevels; ++l) {
      std::vector<int32_t> tb, st(1, 0), li;
      int cur = -1;
      for (auto &pr : per_level[l]) {
        if (pr.first != cur) {
          if (cur >= 0) st.push_back((int32_t)li.size());
          cur = pr.first;
          tb.push_back(cur);
        }
        li.push_back(pr.second);
      }
      if (cur >= 0) st.push_back((int32_t)li.size());
      I["target_boxes_sep_smaller_by_source_level_" + std::to_string(l)] = tb;
      I["from_sep_smaller_by_level_starts_" + std::to_string(l)] = st;
      I["from_sep_smaller_by_level_lists_" + std::to_string(l)] = li;
    }
  }
  // }}}

  // {{{ list 4: coarser leaf sources adjacent to the parent, separated from the box
  {
    std::vector<int32_t> l4s(1, 0), l4l;
    std::unordered_map<int, std::vector<int>> cand_cache;
    std::vector<int> stack;
    for (int b : totp_boxes) {
      const Box &bx = B.boxes[b];
      std::vector<int> found;
      if (bx.level >= 1) {
        int p = bx.parent;
        auto it = cand_cache.find(p);
        if (it == cand_cache.end()) {
          std::vector<int> cand;
          stack.assign(1, 0);
          int lp = B.boxes[p].level;
          while (!stack.empty()) {
            int x = stack.back(); stack.pop_back();
            if (x == p || !B.adjacent(x, p)) continue;
            if (B.boxes[x].leaf) { if (has_own_src(x)) cand.push_back(x); continue; }
            if (B.boxes[x].level >= lp || !has_src(x)) continue;
            for (int k = 0; k < nch; ++k) if (B.boxes[x].child[k]) stack.push_back(B.boxes[x].child[k]);
          }
          it = cand_cache.emplace(p, std::move(cand)).first;
        }
        for (int s : it->second) if (!B.adjacent(s, b)) found.push_back(s);
      }
      std::sort(found.begin(), found.end());
      l4l.insert(l4l.end(), found.begin(), found.end());
      l4s.push_back((int32_t)l4l.size());
    }
    I["from_sep_bigger_starts"] = l4s; I["from_sep_bigger_lists"] = l4l;
  }
  // }}}

  {
    // tree-order particle coordinates
    std::vector<double> ssorted(dim*(size_t)nsrc);
    for (int d = 0; d < dim; ++d)
      for (long i = 0; i < nsrc; ++i) ssorted[(size_t)d*nsrc + i] = src[(size_t)d*nsrc + sperm[i]];
    D["sources"] = ssorted;
    if (!same) {
      std::vector<double> ts(dim*(size_t)ntgt);
      for (int d = 0; d < dim; ++d)
        for (long i = 0; i < ntgt; ++i) ts[(size_t)d*ntgt + i] = T[(size_t)d*ntgt + tperm[i]];
      D["targets"] = ts;
    }
  }
  return res;
}

const char *sumpytree_error(void *h) {
  Result *r = (Result *)h;
  return r->error.empty() ? nullptr : r->error.c_str();
}

void sumpytree_scalars(void *h, int *nlevels, double *root_extent, double *bbox_min) {
  Result *r = (Result *)h;
  *nlevels = r->nlevels;
  *root_extent = r->root_extent;
  for (int d = 0; d < 3; ++d) bbox_min[d] = r->bbox_min[d];
}

// Size of a named array (-1 if absent); copies it into out if out != nullptr.
long sumpytree_get_int(void *h, const char *name, int32_t *out) {
  Result *r = (Result *)h;
  auto it = r->ints.find(name);
  if (it == r->ints.end()) return -1;
  if (out) std::memcpy(out, it->second.data(), it->second.size()*sizeof(int32_t));
  return (long)it->second.size();
}

long sumpytree_get_double(void *h, const char *name, double *out) {
  Result *r = (Result *)h;
  auto it = r->dbls.find(name);
  if (it == r->dbls.end()) return -1;
  if (out) std::memcpy(out, it->second.data(), it->second.size()*sizeof(double));
  return (long)it->second.size();
}

void sumpytree_free(void *h) { delete (Result *)h; }

}  // extern "C"
