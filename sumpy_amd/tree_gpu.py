"""The tree, traversal and translation classes built on the GPU.

Same data model, box numbering, particle order and lists as the host builder
(:func:`sumpy_amd.tree.build_tree_and_traversal`, ``csrc/tree.cpp``), array
for array; the reference takes these from boxtree (``TreeBuilder``,
``FMMTraversalBuilder``, ``TranslationClassesBuilder``:
``sumpy/test/test_fmm.py:182-196``, read at ``sumpy/fmm.py:357-592``).

The steps (kernels in ``csrc/tree_gpu.hip``, scans and compaction here):

1. bounding box (one host synchronisation), Morton keys, a stable sort of the
   source and target keys;
2. boxes level by level: the boxes of a level holding more than
   ``max_particles_in_box`` particles are split; a split box's children are
   the non-empty child cells of its particle ranges (binary search in the
   sorted keys), numbered level by level in Morton order;
3. box arrays (counts, parents, children, centres, flags) and box classes;
4. colleagues, lists 1-4 (count pass, scan, write pass) and the list-2
   translation classes.

Arrays stay on the device; host copies are made once at the end (the
wrangler and the M2L scheduler read some lists on the host).
"""

from __future__ import annotations

import time

import numpy as np
import torch

from sumpy_amd import _capi

_I32 = torch.int32


def _i32(t):
    return t.to(_I32).contiguous()


def _excl(counts):
    """Exclusive scan (int32), with the total appended."""
    out = torch.zeros(counts.numel() + 1, dtype=torch.int64, device=counts.device)
    torch.cumsum(counts.to(torch.int64), 0, out=out[1:])
    return _i32(out)


def _level_starts(levels_of_boxes, nlevels):
    cnt = torch.bincount(levels_of_boxes.to(torch.int64), minlength=nlevels)
    return _excl(cnt[:nlevels])


def _bounds(x):
    """(2, dim): per-axis minimum and maximum of the points *x* (dim, n); in
    two stages over rows of 1024, which keeps the reduction wide."""
    dim, n = x.shape
    m = n//1024*1024
    parts = []
    if m:
        h = x[:, :m].reshape(dim, -1, 1024)
        parts.append(torch.stack([h.amin(2).amin(1), h.amax(2).amax(1)]))
    if m < n:
        t = x[:, m:]
        parts.append(torch.stack([t.amin(1), t.amax(1)]))
    b = parts[0] if len(parts) == 1 else torch.stack(
        [torch.minimum(parts[0][0], parts[1][0]), torch.maximum(parts[0][1], parts[1][1])])
    return b


class _Timer:
    def __init__(self, sync):
        self.sync, self.marks, self.t0 = sync, {}, time.perf_counter()

    def mark(self, name):
        if self.sync:
            torch.cuda.synchronize()
        t = time.perf_counter()
        self.marks[name] = t - self.t0
        self.t0 = t


def build_device(actx, sources, targets=None, max_particles_in_box=30, max_levels=None,
                 timing=None):
    """Build the tree and traversal on the GPU.  Returns ``(sc, tree arrays,
    traversal arrays)`` as device tensors (the host builder's names and
    layouts), plus the per-level list-3 triples under ``"sep_smaller"``.
    *timing*: a dict that receives the seconds per step (synchronising)."""
    dev = actx.device
    st = actx.stream_handle
    tm = _Timer(timing is not None)
    src = torch.as_tensor(sources, dtype=torch.float64, device=dev).contiguous()
    same = targets is None
    tgt = src if same else torch.as_tensor(targets, dtype=torch.float64,
                                           device=dev).contiguous()
    dim, nsrc = src.shape
    ntgt = tgt.shape[1]
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    if nsrc + (0 if same else ntgt) == 0:
        raise ValueError("no particles")
    if max_particles_in_box < 1:
        raise ValueError("max_particles_in_box must be >= 1")
    nch = 1 << dim

    # {{{ bounding box and keys
    lohi = torch.stack([_bounds(src)] if same else [_bounds(src), _bounds(tgt)])
    lohi = torch.stack([lohi[:, 0].amin(0), lohi[:, 1].amax(0)]).cpu().numpy()
    lo, hi = lohi[0], lohi[1]
    ext = 0.0
    for d in range(dim):
        ext = max(ext, float(hi[d] - lo[d]))
    if ext == 0:
        ext = 1.0
    ext *= 1 + 1e-4
    maxdepth = min(max_levels if max_levels and max_levels > 0 else 21, 21 if dim == 3 else 31)
    lo3 = np.zeros(3)
    lo3[:dim] = lo

    def keys_of(x):
        k = torch.empty(x.shape[1], dtype=torch.int64, device=dev)
        _capi.call("sumpytree_gpu_morton", dim, x.shape[1], _capi.ptr(x[0]), _capi.ptr(x[1]),
                   _capi.ptr(x[dim - 1]), lo3.ctypes.data, ext, maxdepth, _capi.ptr(k), st)
        return torch.sort(k, stable=True)

    skeys, sperm = keys_of(src)
    if same:
        tkeys, tperm = None, sperm
    else:
        tkeys, tperm = keys_of(tgt)
    tm.mark("keys")
    # }}}

    # {{{ boxes, level by level
    zero = torch.zeros(1, dtype=_I32, device=dev)
    pre = [torch.zeros(1, dtype=torch.int64, device=dev)]
    slo, shi = [zero.clone()], [torch.full((1,), nsrc, dtype=_I32, device=dev)]
    tlo, thi = [zero.clone()], [torch.full((1,), ntgt, dtype=_I32, device=dev)]
    parent = [zero.clone()]
    split_of = []           # per level: bool per box
    child_parent_slot = []  # per level >= 1: (parent box id, child index) per box
    base = 0
    level = 0
    while True:
        cnt = (shi[-1] - slo[-1]) + (0 if same else (thi[-1] - tlo[-1]))
        split = (cnt > max_particles_in_box) & (level + 1 < maxdepth)
        split_of.append(split)
        idx = torch.nonzero(split).flatten()
        nsplit = int(idx.numel())
        if nsplit == 0:
            break
        e = nsplit*nch
        cs = [torch.empty(e, dtype=_I32, device=dev) for _ in range(5)]
        shift = dim*(maxdepth - level - 1)
        # the split boxes' prefixes and ranges (held until the launch is queued)
        sp = [pre[-1][idx].contiguous()] + [a[-1][idx].contiguous() for a in (slo, shi, tlo, thi)]
        _capi.call("sumpytree_gpu_children", dim, nsplit, *[_capi.ptr(a) for a in sp],
                   _capi.ptr(skeys), None if same else _capi.ptr(tkeys), shift,
                   *[_capi.ptr(c) for c in cs], st)
        ex = torch.nonzero(cs[4]).flatten()
        nb = int(ex.numel())
        par_local = idx[ex // nch]
        pre.append(((pre[-1][par_local] << dim) | (ex % nch)).contiguous())
        slo.append(cs[0][ex]); shi.append(cs[1][ex])
        tlo.append(cs[2][ex]); thi.append(cs[3][ex])
        parent.append(_i32(par_local + base))
        child_parent_slot.append((par_local + base, ex % nch))
        base += int(pre[-2].numel())
        level += 1
        if nb == 0:
            break
    nlevels = len(pre)
    level_counts = [int(p.numel()) for p in pre]
    nboxes = sum(level_counts)
    prefix = torch.cat(pre)
    box_levels = _i32(torch.repeat_interleave(
        torch.arange(nlevels, device=dev), torch.tensor(level_counts, device=dev)))
    bss = torch.cat(slo)
    bscc = torch.cat(shi) - bss
    leaf = ~torch.cat(split_of)
    bts = torch.cat(tlo) if not same else bss
    btcc = (torch.cat(thi) - bts) if not same else bscc
    bscn = torch.where(leaf, bscc, 0)
    btcn = torch.where(leaf, btcc, 0)
    child = torch.zeros(nch*nboxes, dtype=_I32, device=dev)
    start = level_counts[0]
    for lev, (pid, q) in enumerate(child_parent_slot):
        n = level_counts[lev + 1]
        child[q*nboxes + pid] = torch.arange(start, start + n, dtype=_I32, device=dev)
        start += n
    child = child.view(nch, nboxes)
    parent_ids = torch.cat(parent)
    level_start = torch.tensor(np.concatenate([[0], np.cumsum(level_counts)]), dtype=_I32,
                               device=dev)
    centers = torch.empty((dim, nboxes), dtype=torch.float64, device=dev)
    ix = torch.empty((dim, nboxes), dtype=_I32, device=dev)
    _capi.call("sumpytree_gpu_geometry", dim, nboxes, _capi.ptr(prefix), _capi.ptr(box_levels),
               lo3.ctypes.data, ext, _capi.ptr(centers), _capi.ptr(ix), st)
    flags = (_i32(bscn > 0) | (_i32(btcn > 0) << 1) | (_i32(~leaf & (bscc > 0)) << 2)
             | (_i32(~leaf & (btcc > 0)) << 3))
    tm.mark("boxes")
    # }}}

    th = {
        "level_start_box_nrs": level_start, "box_source_starts": _i32(bss),
        "box_source_counts_cumul": _i32(bscc), "box_source_counts_nonchild": _i32(bscn),
        "box_target_starts": _i32(bts), "box_target_counts_cumul": _i32(btcc),
        "box_target_counts_nonchild": _i32(btcn), "box_parent_ids": parent_ids,
        "box_levels": box_levels, "box_flags": _i32(flags),
        "user_source_ids": _i32(sperm),
        "box_child_ids": child, "box_centers": centers,
    }
    sti = torch.empty(ntgt, dtype=_I32, device=dev)
    sti[tperm] = torch.arange(ntgt, dtype=_I32, device=dev)
    th["sorted_target_ids"] = sti
    th["sources"] = src[:, sperm].contiguous()
    th["targets"] = th["sources"] if same else tgt[:, tperm].contiguous()

    # {{{ box classes
    vh = {}
    for name, mask in (("source_boxes", bscn > 0), ("source_parent_boxes", ~leaf & (bscc > 0)),
                       ("target_boxes", btcn > 0), ("target_or_target_parent_boxes", btcc > 0)):
        boxes = _i32(torch.nonzero(mask).flatten())
        vh[name] = boxes
        vh["level_start_" + name[:-2] + "_nrs"] = _level_starts(box_levels[boxes], nlevels)
    # }}}

    # {{{ colleagues and lists
    leaf32 = _i32(leaf)
    view = (dim, nboxes, nlevels, _capi.ptr(box_levels), _capi.ptr(ix), _capi.ptr(child),
            _capi.ptr(parent_ids), _capi.ptr(leaf32), _capi.ptr(level_start),
            _capi.ptr(th["box_source_counts_cumul"]), _capi.ptr(th["box_source_counts_nonchild"]),
            _capi.ptr(prefix))
    ncol = 26 if dim == 3 else 8
    coll = torch.empty(nboxes*ncol, dtype=_I32, device=dev)
    _capi.call("sumpytree_gpu_colleagues", *view, _capi.ptr(coll), st)
    tm.mark("colleagues")

    def two_pass(which, boxes, per_box=1):
        nb = int(boxes.numel())
        counts = torch.zeros(nb*per_box, dtype=_I32, device=dev)
        _capi.call("sumpytree_gpu_list", which, *view, _capi.ptr(coll), nb, _capi.ptr(boxes),
                   _capi.ptr(counts), None, None, None, None, st)
        return counts

    tb, totp = vh["target_boxes"], vh["target_or_target_parent_boxes"]
    # list 1
    c1 = two_pass(1, tb)
    s1 = _excl(c1)
    l1 = torch.empty(int(s1[-1]), dtype=_I32, device=dev)
    if l1.numel():
        _capi.call("sumpytree_gpu_list", 1, *view, _capi.ptr(coll), int(tb.numel()),
                   _capi.ptr(tb), None, _capi.ptr(s1), _capi.ptr(l1), None, None, st)
    vh["neighbor_source_boxes_starts"], vh["neighbor_source_boxes_lists"] = s1, l1
    # list 2 and its translation classes
    nslots = 343 if dim == 3 else 49
    c2 = two_pass(2, totp)
    s2 = _excl(c2)
    n2 = int(s2[-1])
    l2 = torch.empty(n2, dtype=_I32, device=dev)
    slot = torch.empty(n2, dtype=_I32, device=dev)
    used = torch.zeros(nlevels*nslots, dtype=_I32, device=dev)
    if n2:
        _capi.call("sumpytree_gpu_list", 2, *view, _capi.ptr(coll), int(totp.numel()),
                   _capi.ptr(totp), None, _capi.ptr(s2), _capi.ptr(l2), _capi.ptr(slot),
                   _capi.ptr(used), st)
    used_h = used.view(nlevels, nslots).cpu().numpy() > 0
    slot_to_cls = np.full((nlevels, nslots), -1, np.int32)
    gid = 0
    cls_ls = np.zeros(nlevels + 1, np.int32)
    vecs = []
    for lev in range(nlevels):
        cls_ls[lev] = gid
        h = ext/float(1 << lev)
        for s_ in np.nonzero(used_h[lev])[0]:
            slot_to_cls[lev, s_] = gid
            gid += 1
            rem, v = int(s_), [0]*dim
            for d in range(dim - 1, -1, -1):
                v[d] = rem % 7 - 3
                rem //= 7
            vecs.append([v[d]*h for d in range(dim)])
    cls_ls[nlevels] = gid
    entry_level = _i32(torch.repeat_interleave(box_levels[totp], c2.to(torch.int64)))
    cls = torch.empty(n2, dtype=_I32, device=dev)
    s2c = torch.from_numpy(slot_to_cls.ravel()).to(dev)
    _capi.call("sumpytree_gpu_list2_classes", n2, _capi.ptr(entry_level), _capi.ptr(slot),
               _capi.ptr(s2c), nslots, _capi.ptr(cls), st)
    vh["from_sep_siblings_starts"], vh["from_sep_siblings_lists"] = s2, l2
    vh["from_sep_siblings_translation_classes"] = cls
    vh["from_sep_siblings_translation_classes_level_starts"] = torch.from_numpy(cls_ls).to(dev)
    vh["from_sep_siblings_translation_class_to_distance_vector"] = torch.tensor(
        np.array(vecs, np.float64).reshape(-1, dim).T.copy(), device=dev)
    # list 3, per source level
    nt = int(tb.numel())
    c3 = two_pass(3, tb, nlevels).view(nt, nlevels)
    off = _excl(c3.t().reshape(-1))
    l3 = torch.empty(int(off[-1]), dtype=_I32, device=dev)
    if l3.numel():
        _capi.call("sumpytree_gpu_list", 3, *view, _capi.ptr(coll), nt, _capi.ptr(tb), None,
                   _capi.ptr(off), _capi.ptr(l3), None, None, st)
    sep = []
    off_h = off.cpu().numpy()
    for lev in range(nlevels):
        cl = c3[:, lev]
        sel = torch.nonzero(cl).flatten()
        a, b = int(off_h[lev*nt]), int(off_h[(lev + 1)*nt])
        sep.append((_i32(tb[sel]), _excl(cl[sel]), l3[a:b]))
    vh["sep_smaller"] = sep
    # list 4
    c4 = two_pass(4, totp)
    s4 = _excl(c4)
    l4 = torch.empty(int(s4[-1]), dtype=_I32, device=dev)
    if l4.numel():
        _capi.call("sumpytree_gpu_list", 4, *view, _capi.ptr(coll), int(totp.numel()),
                   _capi.ptr(totp), None, _capi.ptr(s4), _capi.ptr(l4), None, None, st)
    vh["from_sep_bigger_starts"], vh["from_sep_bigger_lists"] = s4, l4
    tm.mark("lists")
    # }}}

    sc = dict(dim=dim, nlevels=nlevels, root_extent=ext, bb_min=np.asarray(lo, np.float64),
              same=same)
    if timing is not None:
        timing.update(tm.marks)
    return sc, th, vh


def build_tree_and_traversal_gpu(actx, sources, targets=None, max_particles_in_box=30,
                                 max_levels=None, timing=None):
    """:func:`sumpy_amd.tree.build_tree_and_traversal` with the building on
    the GPU; the same tree, traversal and translation classes."""
    from sumpy_amd.tree import _wrap_device
    sc, th, vh = build_device(actx, sources, targets, max_particles_in_box, max_levels, timing)
    t0 = time.perf_counter()
    tree, trav = _wrap_device(actx, sc, th, vh)
    tree.__dict__["_build_params"] = (max_particles_in_box, max_levels)
    tree.__dict__["_built_on_device"] = True
    if timing is not None:
        timing["host_copies"] = time.perf_counter() - t0
    return tree, trav


__all__ = ["build_device", "build_tree_and_traversal_gpu"]
