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

    @property
    def nsources(self):
        return self.host["sources"].shape[1]

    @property
    def ntargets(self):
        return self.host["targets"].shape[1]

    @property
    def nboxes(self):
        return len(self.host["box_levels"])

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


def build_tree_and_traversal(actx, sources, targets=None, max_particles_in_box=30,
                             max_levels=None):
    """Builds tree + traversal + translation classes in one host pass."""
    src = np.ascontiguousarray(_to_host_coords(sources))
    dim, nsrc = src.shape
    tgt = None if targets is None else np.ascontiguousarray(_to_host_coords(targets))
    lib = _treelib()
    h = lib.sumpytree_build(
        dim, nsrc, src.ctypes.data, 0 if tgt is None else tgt.shape[1],
        None if tgt is None else tgt.ctypes.data, int(max_particles_in_box),
        int(max_levels or 0))
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
    finally:
        lib.sumpytree_free(h)

    tree = Tree(dimensions=dim, nlevels=nlevels, root_extent=root_extent,
                bounding_box=(bb_min, bb_min + root_extent),
                sources_are_targets=tgt is None, host=th)
    dev = {k: actx.from_numpy(v) for k, v in th.items()}
    if tgt is None:
        dev["targets"] = dev["sources"]
    tree.__dict__["_dev"] = dev

    trav = FMMTraversalInfo(tree=tree, host=vh)
    tdev = {k: actx.from_numpy(v) for k, v in vh.items()
            if not k.startswith("from_sep_siblings_translation")}
    tdev["target_boxes_sep_smaller_by_source_level"] = [
        actx.from_numpy(t) for t, _, _ in sep_smaller]
    tdev["from_sep_smaller_by_level"] = [
        _LevelCSR(actx.from_numpy(s), actx.from_numpy(li)) for _, s, li in sep_smaller]
    trav.__dict__["_dev"] = tdev
    trav.host["sep_smaller"] = sep_smaller

    tcd = TranslationClassesData(
        from_sep_siblings_translation_classes=actx.from_numpy(
            vh["from_sep_siblings_translation_classes"]),
        from_sep_siblings_translation_class_to_distance_vector=actx.from_numpy(
            vh["from_sep_siblings_translation_class_to_distance_vector"]),
        from_sep_siblings_translation_classes_level_starts=actx.from_numpy(
            vh["from_sep_siblings_translation_classes_level_starts"]))
    trav.__dict__["_translation_classes"] = tcd
    return tree, trav


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
