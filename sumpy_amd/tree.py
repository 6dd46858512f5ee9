"""boxtree-compatible tree, traversal and translation-class data.

The reference obtains these from boxtree (``TreeBuilder``,
``FMMTraversalBuilder`` and ``TranslationClassesBuilder``, used at
``sumpy/test/test_fmm.py:190-196`` and ``sumpy/fmm.py:357-363``).  boxtree is
not part of this build, so :mod:`sumpy_amd.csrc.tree` (host C++) builds the
same data model and this module exposes it with boxtree's attribute names --
the ones ``SumpyExpansionWrangler`` and ``drive_fmm`` read
(``sumpy/fmm.py:372-592,752-796``).  Arrays are device tensors of the array
context; level starts are also kept on the host.
"""

from __future__ import annotations

import ctypes
import os
from collections.abc import MutableMapping
from dataclasses import dataclass, field

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _treelib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libsumpytree.so")
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is not built; run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.sumpytree_build.restype = ctypes.c_void_p
        lib.sumpytree_build.argtypes = [
            ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_long,
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.sumpytree_build_in_box.restype = ctypes.c_void_p
        lib.sumpytree_build_in_box.argtypes = [
            ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_long,
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double]
        lib.sumpytree_error.restype = ctypes.c_char_p
        lib.sumpytree_error.argtypes = [ctypes.c_void_p]
        lib.sumpytree_scalars.argtypes = [
            ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.sumpytree_get_int.restype = ctypes.c_long
        lib.sumpytree_get_int.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                          ctypes.c_void_p]
        lib.sumpytree_get_double.restype = ctypes.c_long
        lib.sumpytree_get_double.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                             ctypes.c_void_p]
        lib.sumpytree_free.argtypes = [ctypes.c_void_p]
        _lib = lib
    return _lib


def _get_int(h, name):
    lib = _treelib()
    n = lib.sumpytree_get_int(h, name.encode(), None)
    if n < 0:
        raise KeyError(name)
    out = np.empty(n, np.int32)
    lib.sumpytree_get_int(h, name.encode(), out.ctypes.data)
    return out


def _get_double(h, name):
    lib = _treelib()
    n = lib.sumpytree_get_double(h, name.encode(), None)
    if n < 0:
        raise KeyError(name)
    out = np.empty(n, np.float64)
    lib.sumpytree_get_double(h, name.encode(), out.ctypes.data)
    return out


def _to_host_coords(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy().astype(np.float64)
    if isinstance(x, (list, tuple)) or (isinstance(x, np.ndarray) and x.dtype == object):
        return np.stack([_to_host_coords(c) for c in x])
    return np.asarray(x, dtype=np.float64)


@dataclass
class Tree:
    """Tree data (boxtree ``Tree`` attribute names)."""

    dimensions: int
    nlevels: int
    root_extent: float
    bounding_box: tuple
    sources_are_targets: bool
    stick_out_factor: float = 0.0
    host: dict = field(default_factory=dict)

    def __getattr__(self, name):
        dev = self.__dict__.get("_dev")
        if dev is not None and name in dev:
            return dev[name]
        raise AttributeError(name)

    def _shape(self, name):
        dev = self.__dict__.get("_dev")
        return tuple((dev if dev is not None and name in dev else self.host)[name].shape)

    @property
    def nsources(self):
        return self._shape("sources")[1]

    @property
    def ntargets(self):
        return self._shape("targets")[1]

    @property
    def nboxes(self):
        return self._shape("box_levels")[0]

    @property
    def aligned_nboxes(self):
        return self.nboxes

    @property
    def level_start_box_nrs_host(self):
        return self.host["level_start_box_nrs"]


class _LevelCSR:
    def __init__(self, starts, lists):
        self.starts = starts
        self.lists = lists


@dataclass
class FMMTraversalInfo:
    """Interaction lists (boxtree ``FMMTraversalInfo`` attribute names)."""

    tree: Tree
    host: dict = field(default_factory=dict)

    def __getattr__(self, name):
        dev = self.__dict__.get("_dev")
        if dev is not None and name in dev:
            return dev[name]
        raise AttributeError(name)

    # close-interaction lists only exist for particles with extent
    from_sep_close_smaller_starts = None
    from_sep_close_smaller_lists = None
    from_sep_close_bigger_starts = None
    from_sep_close_bigger_lists = None


@dataclass
class TranslationClassesData:
    """boxtree ``TranslationClassesInfo`` fields used by the wrangler."""

    from_sep_siblings_translation_classes: torch.Tensor
    from_sep_siblings_translation_class_to_distance_vector: torch.Tensor
    from_sep_siblings_translation_classes_level_starts: torch.Tensor


_TREE_INT_FIELDS = [
    "level_start_box_nrs", "box_source_starts", "box_source_counts_cumul",
    "box_source_counts_nonchild", "box_target_starts", "box_target_counts_cumul",
    "box_target_counts_nonchild", "box_parent_ids", "box_levels", "box_flags",
    "user_source_ids", "sorted_target_ids"]

_TRAV_INT_FIELDS = [
    "source_boxes", "level_start_source_box_nrs", "source_parent_boxes",
    "level_start_source_parent_box_nrs", "target_boxes",
    "level_start_target_box_nrs", "target_or_target_parent_boxes",
    "level_start_target_or_target_parent_box_nrs",
    "neighbor_source_boxes_starts", "neighbor_source_boxes_lists",
    "from_sep_siblings_starts", "from_sep_siblings_lists",
    "from_sep_bigger_starts", "from_sep_bigger_lists",
    "from_sep_siblings_translation_classes",
    "from_sep_siblings_translation_classes_level_starts"]


def _build_host(sources, targets=None, max_particles_in_box=30, max_levels=None,
                root_box=None):
    """One host pass of the C++ builder: (scalars, tree arrays, traversal
    arrays).  *root_box* = (box_min, extent) fixes the root box."""
    src = np.ascontiguousarray(_to_host_coords(sources))
    dim, nsrc = src.shape
    tgt = None if targets is None else np.ascontiguousarray(_to_host_coords(targets))
    lib = _treelib()
    common = (dim, nsrc, src.ctypes.data, 0 if tgt is None else tgt.shape[1],
              None if tgt is None else tgt.ctypes.data, int(max_particles_in_box),
              int(max_levels or 0))
    if root_box is None:
        h = lib.sumpytree_build(*common)
    else:
        bmin = np.ascontiguousarray(root_box[0], np.float64)
        h = lib.sumpytree_build_in_box(*common, bmin.ctypes.data, float(root_box[1]))
    try:
        err = lib.sumpytree_error(h)
        if err:
            raise ValueError(err.decode())
        nlev = ctypes.c_int()
        ext = ctypes.c_double()
        bmin = (ctypes.c_double * 3)()
        lib.sumpytree_scalars(h, ctypes.byref(nlev), ctypes.byref(ext), bmin)
        nlevels, root_extent = nlev.value, ext.value
        bb_min = np.array(bmin[:dim])
        th = {n: _get_int(h, n) for n in _TREE_INT_FIELDS}
        nboxes = len(th["box_levels"])
        th["box_child_ids"] = _get_int(h, "box_child_ids").reshape(2**dim, nboxes)
        th["box_centers"] = _get_double(h, "box_centers").reshape(dim, nboxes)
        th["sources"] = _get_double(h, "sources").reshape(dim, nsrc)
        th["targets"] = (th["sources"] if tgt is None
                         else _get_double(h, "targets").reshape(dim, tgt.shape[1]))
        vh = {n: _get_int(h, n) for n in _TRAV_INT_FIELDS}
        vecs = _get_double(h, "from_sep_siblings_translation_class_to_distance_vector")
        vh["from_sep_siblings_translation_class_to_distance_vector"] = \
            vecs.reshape(dim, -1)
        sep_smaller = []
        for lev in range(nlevels):
            sep_smaller.append((
                _get_int(h, f"target_boxes_sep_smaller_by_source_level_{lev}"),
                _get_int(h, f"from_sep_smaller_by_level_starts_{lev}"),
                _get_int(h, f"from_sep_smaller_by_level_lists_{lev}")))
        vh["sep_smaller"] = sep_smaller
    finally:
        lib.sumpytree_free(h)
    return dict(dim=dim, nlevels=nlevels, root_extent=root_extent, bb_min=bb_min,
                same=tgt is None), th, vh


def _wrap(actx, sc, th, vh):
    """Tree / traversal / translation-class objects over the host arrays."""
    dim, nlevels = sc["dim"], sc["nlevels"]
    tree = Tree(dimensions=dim, nlevels=nlevels, root_extent=sc["root_extent"],
                bounding_box=(sc["bb_min"], sc["bb_min"] + sc["root_extent"]),
                sources_are_targets=sc["same"], host=th)
    dev = {k: actx.from_numpy(v) for k, v in th.items()}
    if sc["same"]:
        dev["targets"] = dev["sources"]
    tree.__dict__["_dev"] = dev

    sep_smaller = vh["sep_smaller"]
    trav = FMMTraversalInfo(tree=tree, host=vh)
    tdev = {k: actx.from_numpy(v) for k, v in vh.items()
            if not k.startswith("from_sep_siblings_translation") and k != "sep_smaller"}
    tdev["target_boxes_sep_smaller_by_source_level"] = [
        actx.from_numpy(t) for t, _, _ in sep_smaller]
    tdev["from_sep_smaller_by_level"] = [
        _LevelCSR(actx.from_numpy(s), actx.from_numpy(li)) for _, s, li in sep_smaller]
    trav.__dict__["_dev"] = tdev

    tcd = TranslationClassesData(
        from_sep_siblings_translation_classes=actx.from_numpy(
            vh["from_sep_siblings_translation_classes"]),
        from_sep_siblings_translation_class_to_distance_vector=actx.from_numpy(
            vh["from_sep_siblings_translation_class_to_distance_vector"]),
        from_sep_siblings_translation_classes_level_starts=actx.from_numpy(
            vh["from_sep_siblings_translation_classes_level_starts"]))
    trav.__dict__["_translation_classes"] = tcd
    return tree, trav


class _HostCopies(MutableMapping):
    """Host (numpy) copies of device arrays, each made on its first access
    (a device synchronisation) and kept: a tree built on the GPU hands the
    host only what its users read.  ``aliases`` maps a name to another
    whose copy it shares; ``"sep_smaller"`` holds per-level triples."""

    def __init__(self, dev, aliases=None):
        self._dev = dict(dev)
        self._aliases = dict(aliases or {})
        self._copies = {}

    def __getitem__(self, name):
        got = self._copies.get(name)
        if got is None:
            if name in self._aliases:
                got = self[self._aliases[name]]
            else:
                v = self._dev[name]
                got = ([tuple(a.cpu().numpy() for a in trip) for trip in v]
                       if name == "sep_smaller" else v.cpu().numpy())
            self._copies[name] = got
        return got

    def __setitem__(self, name, value):
        self._dev.pop(name, None)
        self._aliases.pop(name, None)
        self._copies[name] = value

    def __delitem__(self, name):
        found = False
        for d in (self._dev, self._aliases, self._copies):
            if name in d:
                del d[name]
                found = True
        if not found:
            raise KeyError(name)

    def _names(self):
        return list(dict.fromkeys([*self._dev, *self._aliases, *self._copies]))

    def __iter__(self):
        return iter(self._names())

    def __len__(self):
        return len(self._names())

    def __contains__(self, name):
        return name in self._dev or name in self._aliases or name in self._copies


def _wrap_device(actx, sc, th_dev, vh_dev):
    """Tree / traversal / translation-class objects over device arrays
    (built on the GPU); host copies are made on first access."""
    th = _HostCopies({k: v for k, v in th_dev.items()
                      if not (sc["same"] and k == "targets")},
                     {"targets": "sources"} if sc["same"] else None)
    vh = _HostCopies(vh_dev)
    dim, nlevels = sc["dim"], sc["nlevels"]
    tree = Tree(dimensions=dim, nlevels=nlevels, root_extent=sc["root_extent"],
                bounding_box=(sc["bb_min"], sc["bb_min"] + sc["root_extent"]),
                sources_are_targets=sc["same"], host=th)
    tree.__dict__["_dev"] = dict(th_dev)
    trav = FMMTraversalInfo(tree=tree, host=vh)
    tdev = {k: v for k, v in vh_dev.items()
            if not k.startswith("from_sep_siblings_translation") and k != "sep_smaller"}
    tdev["target_boxes_sep_smaller_by_source_level"] = [t for t, _, _ in vh_dev["sep_smaller"]]
    tdev["from_sep_smaller_by_level"] = [_LevelCSR(s, li) for _, s, li in vh_dev["sep_smaller"]]
    trav.__dict__["_dev"] = tdev
    trav.__dict__["_translation_classes"] = TranslationClassesData(
        from_sep_siblings_translation_classes=vh_dev["from_sep_siblings_translation_classes"],
        from_sep_siblings_translation_class_to_distance_vector=vh_dev[
            "from_sep_siblings_translation_class_to_distance_vector"],
        from_sep_siblings_translation_classes_level_starts=vh_dev[
            "from_sep_siblings_translation_classes_level_starts"])
    return tree, trav


def build_tree_and_traversal(actx, sources, targets=None, max_particles_in_box=30,
                             max_levels=None):
    """Builds tree + traversal + translation classes in one host pass."""
    sc, th, vh = _build_host(sources, targets, max_particles_in_box, max_levels)
    tree, trav = _wrap(actx, sc, th, vh)
    tree.__dict__["_build_params"] = (max_particles_in_box, max_levels)
    return tree, trav


def user_order_points(tree, which):
    """The tree's ``"sources"`` or ``"targets"`` in the caller's order, on
    the tree's device (``(dim, n)``)."""
    pts = getattr(tree, which)
    if which == "sources":
        ids = tree.host["user_source_ids"]
    else:
        # sorted_target_ids[u] is the tree position of user target u
        return pts[:, torch.from_numpy(tree.host["sorted_target_ids"].astype(np.int64))
                   .to(pts.device)]
    out = torch.empty_like(pts)
    out[:, torch.from_numpy(ids.astype(np.int64)).to(pts.device)] = pts
    return out


def swapped_traversal(actx, traversal):
    """The traversal of the same boxes with the roles of sources and targets
    exchanged: its targets are this tree's sources and its sources this
    tree's targets.  The builder splits on the joint particle count and
    numbers boxes by position only, so the swapped tree has the same boxes
    (checked); every interaction list of the swapped traversal is then the
    transpose of this one's (list 1 of the swapped target box b holds the
    boxes whose list 1 holds b, list 2 likewise, lists 3 and 4 exchange).
    Held on the traversal after the first call."""
    tree = traversal.tree
    if tree.sources_are_targets:
        return traversal
    got = traversal.__dict__.get("_swapped")
    if got is not None:
        return got
    params = tree.__dict__.get("_build_params")
    if params is None:
        raise NotImplementedError("swapped_traversal: the tree was not built by "
                                  "build_tree_and_traversal (a forest?)")
    src = user_order_points(tree, "sources")
    tgt = user_order_points(tree, "targets")
    if tree.__dict__.get("_built_on_device"):
        from sumpy_amd.tree_gpu import build_tree_and_traversal_gpu
        stree, strav = build_tree_and_traversal_gpu(actx, tgt, src, *params)
    else:
        stree, strav = build_tree_and_traversal(actx, tgt.cpu().numpy(), src.cpu().numpy(),
                                                *params)
    for name in ("box_levels", "box_parent_ids", "box_child_ids", "level_start_box_nrs"):
        if not np.array_equal(stree.host[name], tree.host[name]):
            raise RuntimeError(f"swapped_traversal: boxes differ ({name})")
    if not np.array_equal(stree.host["box_centers"], tree.host["box_centers"]) or \
            stree.root_extent != tree.root_extent:
        raise RuntimeError("swapped_traversal: box geometry differs")
    traversal.__dict__["_swapped"] = strav
    return strav


# {{{ forests: many independent problems in one tree

def build_forest(actx, sources_list, max_particles_in_box=30, max_levels=None):
    """One tree and traversal over several independent problems (sources =
    targets in each), so that every FMM stage processes all of them in one
    launch.

    Each problem gets its own octree, all in one common root box (the union
    of the problems' bounding boxes): box sizes, expansion scales and the
    translation-class vectors then agree level by level.  The trees are merged
    level-major -- the boxes of level l of problem 0, then of problem 1, ... --
    and every list of the traversal is the concatenation of the problems'
    lists in that order, with box numbers mapped; no list couples two
    problems.  Particles are concatenated problem by problem: user index
    ``offsets[i] + j`` is particle ``j`` of problem ``i``.

    Returns ``(tree, traversal, offsets)``."""
    pts = [np.ascontiguousarray(_to_host_coords(s)) for s in sources_list]
    if not pts:
        raise ValueError("no problems")
    dim = pts[0].shape[0]
    lo = np.min([p.min(axis=1) for p in pts if p.shape[1]], axis=0)
    hi = np.max([p.max(axis=1) for p in pts if p.shape[1]], axis=0)
    ext = float((hi - lo).max())*(1 + 1e-4) or 1.0
    parts = [_build_host(p, None, max_particles_in_box, max_levels, (lo, ext))
             for p in pts]
    sc, th, vh = _merge_forest(parts, dim)
    sc["bb_min"] = lo
    tree, trav = _wrap(actx, sc, th, vh)
    offsets = np.concatenate([[0], np.cumsum([p.shape[1] for p in pts])]).astype(np.int64)
    return tree, trav, offsets


def _merge_forest(parts, dim):
    nlev = max(sc["nlevels"] for sc, _, _ in parts)
    np_ = len(parts)
    lbs = [np.concatenate([th["level_start_box_nrs"],
                           np.full(nlev + 1 - len(th["level_start_box_nrs"]),
                                   th["level_start_box_nrs"][-1])]) for _, th, _ in parts]
    # new box number: level-major, then problem, then the old order
    cnt = np.array([[lb[l + 1] - lb[l] for l in range(nlev)] for lb in lbs])  # (np, nlev)
    lstart = np.concatenate([[0], np.cumsum(cnt.sum(axis=0))])
    first = lstart[:-1][None, :] + np.concatenate(
        [np.zeros((1, nlev), np.int64), np.cumsum(cnt, axis=0)[:-1]], axis=0)
    maps = []
    for i, (_, th, _) in enumerate(parts):
        lev = th["box_levels"].astype(np.int64)
        old = np.arange(len(lev))
        maps.append((first[i, lev] + old - lbs[i][lev]).astype(np.int64))
    nboxes = int(lstart[-1])
    nsrc = [th["sources"].shape[1] for _, th, _ in parts]
    soff = np.concatenate([[0], np.cumsum(nsrc)]).astype(np.int64)

    def m(i, a):
        return maps[i][np.asarray(a, np.int64)]

    th = {}
    th["level_start_box_nrs"] = lstart.astype(np.int32)
    for name in ("box_source_starts", "box_target_starts"):
        out = np.zeros(nboxes, np.int64)
        for i, (_, t, _) in enumerate(parts):
            out[maps[i]] = t[name] + soff[i]
        th[name] = out.astype(np.int32)
    for name in ("box_source_counts_cumul", "box_source_counts_nonchild",
                 "box_target_counts_cumul", "box_target_counts_nonchild", "box_levels",
                 "box_flags"):
        out = np.zeros(nboxes, np.int64)
        for i, (_, t, _) in enumerate(parts):
            out[maps[i]] = t[name]
        th[name] = out.astype(np.int32)
    par = np.zeros(nboxes, np.int64)
    child = np.zeros((2**dim, nboxes), np.int64)
    centers = np.zeros((dim, nboxes))
    for i, (_, t, _) in enumerate(parts):
        par[maps[i]] = m(i, t["box_parent_ids"])
        c = t["box_child_ids"].astype(np.int64)
        child[:, maps[i]] = np.where(c != 0, maps[i][c], 0)
        centers[:, maps[i]] = t["box_centers"]
    th["box_parent_ids"] = par.astype(np.int32)
    th["box_child_ids"] = child.astype(np.int32)
    th["box_centers"] = centers
    th["user_source_ids"] = np.concatenate(
        [t["user_source_ids"] + soff[i] for i, (_, t, _) in enumerate(parts)]).astype(np.int32)
    th["sorted_target_ids"] = np.concatenate(
        [t["sorted_target_ids"] + soff[i] for i, (_, t, _) in enumerate(parts)]).astype(np.int32)
    th["sources"] = np.concatenate([t["sources"] for _, t, _ in parts], axis=1)
    th["targets"] = th["sources"]

    def level_list(name, starts_name):
        """A box list sorted by level: level by level, problem by problem."""
        out, st = [], [0]
        for lev in range(nlev):
            for i, (_, _, v) in enumerate(parts):
                ls = v[starts_name]
                if lev + 1 < len(ls):
                    out.append(m(i, v[name][ls[lev]:ls[lev + 1]]))
            st.append(sum(len(o) for o in out))
        return np.concatenate(out).astype(np.int32), np.array(st, np.int32)

    def csr_rows(starts_name, lists_name, owner_starts):
        """The CSR rows of a list keyed by a level-sorted box list, in the
        merged order of that box list."""
        out_starts, out_lists, n = [0], [], 0
        for lev in range(nlev):
            for i, (_, _, v) in enumerate(parts):
                ls = v[owner_starts]
                if lev + 1 >= len(ls):
                    continue
                s = v[starts_name]
                r0, r1 = ls[lev], ls[lev + 1]
                seg = v[lists_name][s[r0]:s[r1]]
                out_lists.append(m(i, seg))
                out_starts.extend((s[r0 + 1:r1 + 1] - s[r0] + n).tolist())
                n += int(s[r1] - s[r0])
        lists = np.concatenate(out_lists) if out_lists else np.zeros(0, np.int64)
        return np.array(out_starts, np.int32), lists.astype(np.int32)

    vh = {}
    for name, starts in (("source_boxes", "level_start_source_box_nrs"),
                         ("source_parent_boxes", "level_start_source_parent_box_nrs"),
                         ("target_boxes", "level_start_target_box_nrs"),
                         ("target_or_target_parent_boxes",
                          "level_start_target_or_target_parent_box_nrs")):
        vh[name], vh[starts] = level_list(name, starts)
    vh["neighbor_source_boxes_starts"], vh["neighbor_source_boxes_lists"] = csr_rows(
        "neighbor_source_boxes_starts", "neighbor_source_boxes_lists",
        "level_start_target_box_nrs")
    for pre in ("from_sep_siblings", "from_sep_bigger"):
        vh[pre + "_starts"], vh[pre + "_lists"] = csr_rows(
            pre + "_starts", pre + "_lists", "level_start_target_or_target_parent_box_nrs")
    # translation classes: one class per distinct vector per level, numbered
    # in order of first appearance over the problems
    cls_all, vec_levels, cls_starts = [], [], [0]
    for lev in range(nlev):
        key_to_id, vecs_lev = {}, []
        for i, (sc, _, v) in enumerate(parts):
            ls = v["level_start_target_or_target_parent_box_nrs"]
            if lev + 1 >= len(ls):
                continue
            cs = v["from_sep_siblings_translation_classes_level_starts"]
            vec = v["from_sep_siblings_translation_class_to_distance_vector"]
            h = sc["root_extent"]*2.0**-lev
            s = v["from_sep_siblings_starts"]
            seg = v["from_sep_siblings_translation_classes"][s[ls[lev]]:s[ls[lev + 1]]]
            # this problem's classes of the level -> merged ids (by vector)
            c0 = int(cs[lev]) if lev < len(cs) - 1 else 0
            c1 = int(cs[lev + 1]) if lev < len(cs) - 1 else 0
            lut = np.empty(c1 - c0, np.int64)
            for c in range(c0, c1):
                key = tuple(np.rint(vec[:, c]/h).astype(np.int64).tolist())
                cid = key_to_id.get(key)
                if cid is None:
                    cid = len(key_to_id)
                    key_to_id[key] = cid
                    vecs_lev.append(vec[:, c])
                lut[c - c0] = cid
            cls_all.append(lut[seg.astype(np.int64) - c0] + cls_starts[-1])
        cls_starts.append(cls_starts[-1] + len(key_to_id))
        vec_levels.extend(vecs_lev)
    vh["from_sep_siblings_translation_classes"] = (
        np.concatenate(cls_all) if cls_all else np.zeros(0)).astype(np.int32)
    vh["from_sep_siblings_translation_classes_level_starts"] = np.array(cls_starts, np.int32)
    vh["from_sep_siblings_translation_class_to_distance_vector"] = (
        np.array(vec_levels).T.reshape(dim, -1) if vec_levels else np.zeros((dim, 0)))
    # list 3, by source level
    sep = []
    for lev in range(nlev):
        tb, st, li, n = [], [0], [], 0
        for i, (_, _, v) in enumerate(parts):
            if lev >= len(v["sep_smaller"]):
                continue
            t, s, lst = v["sep_smaller"][lev]
            tb.append(m(i, t))
            li.append(m(i, lst[s[0]:s[-1]]))
            st.extend((s[1:] - s[0] + n).tolist())
            n += int(s[-1] - s[0])
        sep.append((np.concatenate(tb).astype(np.int32) if tb else np.zeros(0, np.int32),
                    np.array(st, np.int32),
                    np.concatenate(li).astype(np.int32) if li else np.zeros(0, np.int32)))
    vh["sep_smaller"] = sep
    sc = dict(dim=dim, nlevels=nlev, root_extent=parts[0][0]["root_extent"], same=True)
    return sc, th, vh

# }}}


class TreeBuilder:
    """``boxtree.TreeBuilder`` call signature: ``tb(actx, sources, targets=None,
    max_particles_in_box=...)`` returning ``(tree, event)``.  The traversal is
    built in the same host pass and handed out by :class:`FMMTraversalBuilder`."""

    def __init__(self, actx=None):
        self.actx = actx

    def __call__(self, actx, sources, targets=None, max_particles_in_box=30,
                 max_levels=None, **kwargs):
        tree, trav = build_tree_and_traversal(
            actx, sources, targets, max_particles_in_box, max_levels)
        tree.__dict__["_traversal"] = trav
        return tree, None


class FMMTraversalBuilder:
    """``boxtree.traversal.FMMTraversalBuilder`` call signature."""

    def __init__(self, actx=None, well_sep_is_n_away=1):
        if well_sep_is_n_away != 1:
            raise NotImplementedError("only well_sep_is_n_away=1 is supported")

    def __call__(self, actx, tree, **kwargs):
        return tree.__dict__["_traversal"], None


class TranslationClassesBuilder:
    """``boxtree.translation_classes.TranslationClassesBuilder`` call
    signature (``sumpy/fmm.py:357-363``)."""

    def __init__(self, actx=None):
        pass

    def __call__(self, actx, trav, tree, is_translation_per_level=True):
        if not is_translation_per_level:
            raise NotImplementedError("translation classes are per level")
        return trav.__dict__["_translation_classes"], None
