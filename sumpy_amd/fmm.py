"""The FMM wrangler: the drop-in boundary above the operators.

:class:`SumpyTreeIndependentDataForWrangler` (operator factory) and
:class:`SumpyExpansionWrangler` keep the constructor arguments, attributes and
method signatures of ``sumpy/fmm.py:99-1115`` (the interface boxtree's
``drive_fmm`` calls), so :func:`sumpy_amd.drive.drive_fmm` -- or boxtree's --
runs them unchanged.  Expansion arrays are flat device tensors with per-level
starts (``build_csr_level_starts``, ``fmm.py:1093-1115``); per-level views are
row-major ``(nboxes_level, ncoeff)`` reshapes.
"""

from __future__ import annotations

from functools import cached_property

import numpy as np
import torch

from sumpy_amd.array_context import obj_array, torch_dtype
from sumpy_amd.e2e import (
    E2EFromChildren,
    E2EFromCSR,
    E2EFromParent,
    M2LGenerateTranslationClassesDependentData,
    M2LPostprocessLocal,
    M2LPreprocessMultipole,
    M2LUsingTranslationClassesDependentData,
)
from sumpy_amd.e2p import E2PFromCSR, E2PFromSingleBox
from sumpy_amd.p2e import P2EFromCSR, P2EFromSingleBox
from sumpy_amd.p2p import P2PFromCSR


def build_csr_level_starts(level_orders, order_to_size, level_starts):
    """Starts of each level's block in a flat per-box array
    (reference ``fmm.py:1093-1115``)."""
    result = [0]
    for lev in range(len(level_orders)):
        lev_nboxes = level_starts[lev + 1] - level_starts[lev]
        result.append(result[-1] + order_to_size(level_orders[lev])*lev_nboxes)
    return result


class SumpyTreeIndependentDataForWrangler:
    """Operator factory (reference ``fmm.py:99-252``).

    Extra keywords: ``p2p_precision``, ``"fp64"`` (plain) or ``"fp32"`` for
    the near field (the M2L precision is chosen by the local expansion's
    ``m2l_translation``, ``VolumeTaylorM2LWithDenseClassMatrices(precision=...)``);
    ``list34_precision``, ``"fp64"`` or ``"fp32"`` for the terms of lists 3
    and 4 (M2P and P2L; FP32 terms, FP64 sums), ``p2m_precision`` for the
    formation of the multipoles (FP32 moments and box sums, FP64 projection)
    and ``l2p_precision`` for the
    evaluation of the local expansions at the targets (potentials), both
    used where all levels share the expansion order;
    ``autotune``, time the M2L contraction's split policies on the first pass
    of each kind of tree and keep the fastest (cached on disk,
    :mod:`sumpy_amd.autotune`).
    """

    def __init__(self, array_context, multipole_expansion_factory,
                 local_expansion_factory, target_kernels, exclude_self=False,
                 use_rscale=None, strength_usage=None, source_kernels=None,
                 p2p_precision="fp64", autotune=False, list34_precision="fp64",
                 l2p_precision="fp64", p2m_precision="fp64"):
        for name, v in (("list34_precision", list34_precision),
                        ("l2p_precision", l2p_precision), ("p2m_precision", p2m_precision)):
            if v not in ("fp64", "fp32"):
                raise ValueError(f"{name}: 'fp64' or 'fp32', got {v!r}")
        self._setup_actx = array_context
        self.multipole_expansion_factory = multipole_expansion_factory
        self.local_expansion_factory = local_expansion_factory
        self.source_kernels = source_kernels
        self.target_kernels = target_kernels
        self.exclude_self = exclude_self
        self.use_rscale = use_rscale
        self.strength_usage = strength_usage
        self.p2p_precision = p2p_precision
        self.list34_precision = list34_precision
        self.l2p_precision = l2p_precision
        self.p2m_precision = p2m_precision
        self.autotune = autotune
        self._memo = {}

    @cached_property
    def m2l_tuner(self):
        from sumpy_amd.autotune import M2LSplitTuner
        return M2LSplitTuner()

    def _m(self, key, make):
        v = self._memo.get(key)
        if v is None:
            v = make()
            self._memo[key] = v
        return v

    def get_base_kernel(self):
        bases = {k.get_base_kernel() for k in self.target_kernels}
        if len(bases) != 1:
            raise ValueError("target kernels must share a base kernel")
        return bases.pop()

    def multipole_expansion(self, order):
        def make():
            if self.use_rscale is None:
                return self.multipole_expansion_factory(order)
            return self.multipole_expansion_factory(order, use_rscale=self.use_rscale)
        return self._m(("mpole", order), make)

    def local_expansion(self, order):
        def make():
            if self.use_rscale is None:
                return self.local_expansion_factory(order)
            return self.local_expansion_factory(order, use_rscale=self.use_rscale)
        return self._m(("local", order), make)

    @property
    def m2l_translation(self):
        return self.local_expansion(0).m2l_translation

    def p2m(self, tgt_order):
        return self._m(("p2m", tgt_order), lambda: P2EFromSingleBox(
            kernels=self.source_kernels, expansion=self.multipole_expansion(tgt_order),
            strength_usage=self.strength_usage, name="p2m"))

    def p2l(self, tgt_order):
        return self._m(("p2l", tgt_order), lambda: P2EFromCSR(
            kernels=self.source_kernels, expansion=self.local_expansion(tgt_order),
            strength_usage=self.strength_usage, name="p2l"))

    def m2m(self, src_order, tgt_order):
        return self._m(("m2m", src_order, tgt_order), lambda: E2EFromChildren(
            self.multipole_expansion(src_order), self.multipole_expansion(tgt_order),
            name="m2m"))

    def m2l(self, src_order, tgt_order, m2l_use_translation_classes_dependent_data=False):
        cls = (M2LUsingTranslationClassesDependentData
               if m2l_use_translation_classes_dependent_data else E2EFromCSR)
        return self._m(("m2l", src_order, tgt_order, cls), lambda: cls(
            self.multipole_expansion(src_order), self.local_expansion(tgt_order),
            name="m2l"))

    def m2l_translation_class_dependent_data_kernel(self, src_order, tgt_order):
        return self._m(("m2l_data", src_order, tgt_order),
                       lambda: M2LGenerateTranslationClassesDependentData(
                           self.multipole_expansion(src_order),
                           self.local_expansion(tgt_order)))

    def m2l_preprocess_mpole_kernel(self, src_order, tgt_order):
        return self._m(("m2l_pre", src_order, tgt_order), lambda: M2LPreprocessMultipole(
            self.multipole_expansion(src_order), self.local_expansion(tgt_order)))

    def m2l_postprocess_local_kernel(self, src_order, tgt_order):
        return self._m(("m2l_post", src_order, tgt_order), lambda: M2LPostprocessLocal(
            self.multipole_expansion(src_order), self.local_expansion(tgt_order)))

    def l2l(self, src_order, tgt_order):
        return self._m(("l2l", src_order, tgt_order), lambda: E2EFromParent(
            self.local_expansion(src_order), self.local_expansion(tgt_order), name="l2l"))

    def m2p(self, src_order):
        return self._m(("m2p", src_order), lambda: E2PFromCSR(
            self.multipole_expansion(src_order), self.target_kernels, name="m2p"))

    def l2p(self, src_order):
        return self._m(("l2p", src_order), lambda: E2PFromSingleBox(
            self.local_expansion(src_order), self.target_kernels, name="l2p"))

    def p2p(self):
        return self._m(("p2p",), lambda: P2PFromCSR(
            target_kernels=self.target_kernels, source_kernels=self.source_kernels,
            exclude_self=self.exclude_self, strength_usage=self.strength_usage,
            name="p2p", precision=self.p2p_precision))


class SumpyExpansionWrangler:
    """Implements boxtree's ``ExpansionWranglerInterface`` with the GPU
    operators (reference ``fmm.py:264-1088``)."""

    def __init__(self, tree_indep, traversal, dtype, fmm_level_to_order,
                 source_extra_kwargs=None, kernel_extra_kwargs=None,
                 self_extra_kwargs=None, translation_classes_data=None,
                 preprocessed_mpole_dtype=None, *, _disable_translation_classes=False):
        self.tree_indep = tree_indep
        self.traversal = traversal
        self.dtype = np.dtype(dtype)
        if not callable(fmm_level_to_order):
            raise TypeError("fmm_level_to_order not passed")
        base_kernel = tree_indep.get_base_kernel()
        self.source_extra_kwargs = dict(source_extra_kwargs or {})
        self.kernel_extra_kwargs = dict(kernel_extra_kwargs or {})
        self.self_extra_kwargs = dict(self_extra_kwargs or {})
        self.extra_kwargs = dict(self.source_extra_kwargs)
        self.extra_kwargs.update(self.kernel_extra_kwargs)
        kernel_arg_set = frozenset(self.kernel_extra_kwargs.items())
        self.level_orders = [fmm_level_to_order(base_kernel, kernel_arg_set, self.tree, lev)
                             for lev in range(self.tree.nlevels)]
        tr = tree_indep.m2l_translation
        self.preprocessed_mpole_dtype = (np.dtype(preprocessed_mpole_dtype)
                                         if preprocessed_mpole_dtype is not None
                                         else tr.preprocessed_dtype(self.dtype))
        if _disable_translation_classes or not base_kernel.is_translation_invariant:
            self.supports_translation_classes = False
        else:
            if translation_classes_data is None:
                from sumpy_amd.tree import TranslationClassesBuilder
                actx = tree_indep._setup_actx
                translation_classes_data, _ = TranslationClassesBuilder(actx)(
                    actx, traversal, self.tree, is_translation_per_level=True)
            self.supports_translation_classes = True
        self.translation_classes_data = translation_classes_data
        #: lists 3 and 4 of all levels in one launch each when the levels
        #: share the expansion order; False: the reference's per-level loop
        self.all_levels_lists = True
        self._memo = {}

    @property
    def tree(self):
        return self.traversal.tree

    def level_to_rscale(self, level):
        """``2 r / order`` with ``r`` the box size (reference ``fmm.py:369-386``)."""
        order = self.level_orders[level]
        r = self.tree.root_extent*(2**-level)
        return r*2/order

    # {{{ data vector utilities

    @cached_property
    def tree_level_start_box_nrs(self):
        return np.asarray(self.tree.host["level_start_box_nrs"])

    def _expansions_level_starts(self, order_to_size):
        return build_csr_level_starts(self.level_orders, order_to_size,
                                      self.tree_level_start_box_nrs)

    @cached_property
    def _mpole_level_starts(self):
        return self._expansions_level_starts(
            lambda order: len(self.tree_indep.multipole_expansion(order)))

    def multipole_expansions_level_starts(self):
        return self._mpole_level_starts

    @cached_property
    def _local_level_starts(self):
        return self._expansions_level_starts(
            lambda order: len(self.tree_indep.local_expansion(order)))

    def local_expansions_level_starts(self):
        return self._local_level_starts

    @cached_property
    def _class_level_starts(self):
        return self.traversal.host["from_sep_siblings_translation_classes_level_starts"]

    def m2l_translation_class_level_start_box_nrs(self):
        return self._class_level_starts

    def m2l_translation_classes_dependent_data_level_starts(self):
        def order_to_size(order):
            mpole = self.tree_indep.multipole_expansion(order)
            local = self.tree_indep.local_expansion(order)
            return local.m2l_translation.translation_classes_dependent_ndata(local, mpole)
        return build_csr_level_starts(self.level_orders, order_to_size,
                                      self.m2l_translation_class_level_start_box_nrs())

    def _zeros(self, n, dtype):
        return torch.zeros(int(n), dtype=torch_dtype(dtype),
                           device=self.tree_indep._setup_actx.device)

    def multipole_expansion_zeros(self, actx):
        return self._zeros(self.multipole_expansions_level_starts()[-1], self.dtype)

    def local_expansion_zeros(self, actx):
        return self._zeros(self.local_expansions_level_starts()[-1], self.dtype)

    def m2l_translation_classes_dependent_data_zeros(self, actx):
        data_starts = self.m2l_translation_classes_dependent_data_level_starts()
        cls_starts = self.m2l_translation_class_level_start_box_nrs()
        result = []
        for lev in range(self.tree.nlevels):
            n = data_starts[lev + 1] - data_starts[lev]
            ncls = cls_starts[lev + 1] - cls_starts[lev]
            result.append(self._zeros(n, np.float64).reshape(ncls, -1)
                          if ncls else self._zeros(0, np.float64).reshape(0, 0))
        return result

    def multipole_expansions_view(self, mpole_exps, level):
        e0, e1 = self.multipole_expansions_level_starts()[level:level + 2]
        b0, b1 = self.tree_level_start_box_nrs[level:level + 2]
        return b0, mpole_exps[e0:e1].reshape(int(b1 - b0), -1)

    def local_expansions_view(self, local_exps, level):
        e0, e1 = self.local_expansions_level_starts()[level:level + 2]
        b0, b1 = self.tree_level_start_box_nrs[level:level + 2]
        return b0, local_exps[e0:e1].reshape(int(b1 - b0), -1)

    def m2l_translation_classes_dependent_data_view(self, data, level):
        cls_start = self.m2l_translation_class_level_start_box_nrs()[level]
        return cls_start, data[level]

    def _m2l_level_sizes(self, nexprs):
        def order_to_size(order):
            mpole = self.tree_indep.multipole_expansion(order)
            local = self.tree_indep.local_expansion(order)
            return nexprs(local.m2l_translation, local, mpole)
        return build_csr_level_starts(self.level_orders, order_to_size,
                                      self.tree_level_start_box_nrs)

    def m2l_preproc_mpole_expansions_level_starts(self):
        return self._m2l_level_sizes(
            lambda tr, loc, mp: tr.preprocess_multipole_nexprs(loc, mp))

    def m2l_preproc_mpole_expansion_zeros(self, actx, template_ary):
        # preprocessing writes every entry of every box: no clearing needed
        starts = self.m2l_preproc_mpole_expansions_level_starts()
        return self._per_level(starts, clear=False)

    def m2l_work_array_level_starts(self):
        return self._m2l_level_sizes(
            lambda tr, loc, mp: tr.postprocess_local_nexprs(loc, mp))

    def m2l_work_array_zeros(self, actx, template_ary):
        return self._per_level(self.m2l_work_array_level_starts(), clear=True)

    def _per_level(self, starts, clear):
        """Per-level views of one allocation (cleared with one fill); the
        allocation and its level starts are kept as ``flat`` and ``starts``."""
        alloc = torch.zeros if clear else torch.empty
        flat = alloc(int(starts[-1]), dtype=torch_dtype(self.preprocessed_mpole_dtype),
                     device=self.tree_indep._setup_actx.device)
        result = _LevelViews()
        for lev in range(self.tree.nlevels):
            b0, b1 = self.tree_level_start_box_nrs[lev:lev + 2]
            result.append(flat[int(starts[lev]):int(starts[lev + 1])].reshape(
                int(b1 - b0), -1))
        result.flat, result.starts = flat, starts
        return result

    def _levels_view(self, flat, starts, l0, l1):
        """Boxes of levels l0..l1 (of one expansion order) as one 2D view."""
        b0, b1 = self.tree_level_start_box_nrs[l0], self.tree_level_start_box_nrs[l1 + 1]
        return flat[int(starts[l0]):int(starts[l1 + 1])].reshape(int(b1 - b0), -1)

    def _cached_zeros(self, name, make):
        """A zero result that is never written, kept for the stages that have
        no work for this traversal (the driver then skips adding it)."""
        z = self._memo.get(name)
        if z is None:
            z = make()
            self._memo[name] = z
        return z

    def m2l_preproc_mpole_expansions_view(self, mpole_exps, level):
        return self.tree_level_start_box_nrs[level], mpole_exps[level]

    m2l_work_array_view = m2l_preproc_mpole_expansions_view

    def output_zeros(self, actx):
        return obj_array([self._zeros(self.tree.ntargets, self.dtype)
                          for _ in self.tree_indep.target_kernels])

    # (the permutations as int64 once, for index_select's gather)
    @cached_property
    def _user_source_ids_long(self):
        return self.tree.user_source_ids.long()

    @cached_property
    def _sorted_target_ids_long(self):
        return self.tree.sorted_target_ids.long()

    def reorder_sources(self, source_array):
        return torch.index_select(source_array, 0, self._user_source_ids_long)

    def reorder_potentials(self, potentials):
        assert isinstance(potentials, np.ndarray) and potentials.dtype.char == "O"
        idx = self._sorted_target_ids_long
        return obj_array([torch.index_select(p, 0, idx) for p in potentials])

    def distribute_source_weights(self, actx, src_weight_vecs, src_idx_all_ranks):
        return src_weight_vecs

    def gather_potential_results(self, actx, potentials, tgt_idx_all_ranks):
        return potentials

    def communicate_mpoles(self, actx, mpole_exps, return_stats=False):
        return None

    @cached_property
    def max_nsources_in_one_box(self):
        return int(self.tree.host["box_source_counts_nonchild"].max())

    @cached_property
    def max_ntargets_in_one_box(self):
        return int(self.tree.host["box_target_counts_nonchild"].max())

    # }}}

    def box_source_list_kwargs(self):
        return {"box_source_starts": self.tree.box_source_starts,
                "box_source_counts_nonchild": self.tree.box_source_counts_nonchild,
                "sources": self.tree.sources}

    def box_target_list_kwargs(self):
        return {"box_target_starts": self.tree.box_target_starts,
                "box_target_counts_nonchild": self.tree.box_target_counts_nonchild,
                "targets": self.tree.targets}

    # {{{ FMM stages (reference fmm.py:609-1086)

    def form_multipoles(self, actx, level_start_source_box_nrs, source_boxes,
                        src_weight_vecs):
        mpoles = self.multipole_expansion_zeros(actx)
        starts = _host(level_start_source_box_nrs)
        kwargs = dict(self.extra_kwargs)
        kwargs.update(self.box_source_list_kwargs())
        if self._all_levels_ok() and source_boxes is getattr(self.traversal, "source_boxes",
                                                              None):
            # one launch for every level's source boxes (same order on all levels)
            p2m = self.tree_indep.p2m(self.level_orders[0])
            p2m(actx, source_boxes=source_boxes, centers=self.tree.box_centers,
                strengths=src_weight_vecs, tgt_expansions=self._one_row_per_box(mpoles),
                tgt_base_ibox=0, rscale=self.level_to_rscale(0),
                box_levels=self.tree.box_levels, level_rscales=self._level_rscales(),
                p2m_precision=self.tree_indep.p2m_precision, **kwargs)
            return mpoles
        for lev in range(self.tree.nlevels):
            start, stop = starts[lev:lev + 2]
            if start == stop:
                continue
            p2m = self.tree_indep.p2m(self.level_orders[lev])
            base, view = self.multipole_expansions_view(mpoles, lev)
            res = p2m(actx, source_boxes=source_boxes[start:stop],
                      centers=self.tree.box_centers, strengths=src_weight_vecs,
                      tgt_expansions=view, tgt_base_ibox=base,
                      rscale=self.level_to_rscale(lev),
                      list34_precision=self.tree_indep.list34_precision, **kwargs)
            assert res is view
        return mpoles

    def coarsen_multipoles(self, actx, level_start_source_parent_box_nrs,
                           source_parent_boxes, mpoles):
        starts = _host(level_start_source_parent_box_nrs)
        for source_level in range(self.tree.nlevels - 1, 2, -1):
            target_level = source_level - 1
            start, stop = starts[target_level:target_level + 2]
            if start == stop:
                continue
            m2m = self.tree_indep.m2m(self.level_orders[source_level],
                                      self.level_orders[target_level])
            sbase, sview = self.multipole_expansions_view(mpoles, source_level)
            tbase, tview = self.multipole_expansions_view(mpoles, target_level)
            res = m2m(actx, src_expansions=sview, src_base_ibox=sbase,
                      tgt_expansions=tview, tgt_base_ibox=tbase,
                      target_boxes=source_parent_boxes[start:stop],
                      box_child_ids=self.tree.box_child_ids,
                      centers=self.tree.box_centers,
                      src_rscale=self.level_to_rscale(source_level),
                      tgt_rscale=self.level_to_rscale(target_level),
                      **self.kernel_extra_kwargs)
            assert res is tview
        return mpoles

    def eval_direct(self, actx, target_boxes, source_box_starts, source_box_lists,
                    src_weight_vecs):
        pot = self.output_zeros(actx)
        kwargs = dict(self.extra_kwargs)
        kwargs.update(self.self_extra_kwargs)
        kwargs.update(self.box_source_list_kwargs())
        kwargs.update(self.box_target_list_kwargs())
        res = self.tree_indep.p2p()(
            actx, target_boxes=target_boxes, source_box_starts=source_box_starts,
            source_box_lists=source_box_lists, strength=src_weight_vecs, result=pot,
            max_nsources_in_one_box=self.max_nsources_in_one_box,
            max_ntargets_in_one_box=self.max_ntargets_in_one_box,
            box_centers=self.tree.box_centers, **kwargs)
        for a, b in zip(pot, res):
            assert a is b
        return pot

    def multipole_to_local_precompute(self):
        memo = self._memo.get("m2l_precompute")
        if memo is not None:
            return memo
        actx = self.tree_indep._setup_actx
        data = self.m2l_translation_classes_dependent_data_zeros(actx)
        result = []
        for lev in range(self.tree.nlevels):
            order = self.level_orders[lev]
            cls_start, view = self.m2l_translation_classes_dependent_data_view(data, lev)
            ncls = view.shape[0]
            if ncls == 0:
                result.append(view)
                continue
            vecs = self.translation_classes_data \
                .from_sep_siblings_translation_class_to_distance_vector
            self.tree_indep.m2l_translation_class_dependent_data_kernel(order, order)(
                actx, src_rscale=self.level_to_rscale(lev),
                translation_classes_level_start=cls_start, ntranslation_classes=ncls,
                m2l_translation_classes_dependent_data=view,
                m2l_translation_vectors=vecs, ntranslation_vectors=vecs.shape[1],
                **self.kernel_extra_kwargs)
            result.append(view)
        self._memo["m2l_precompute"] = result
        return result

    def _add_m2l_precompute_kwargs(self, kwargs, lev):
        if not self.supports_translation_classes:
            return
        data = self.multipole_to_local_precompute()
        cls_start, view = self.m2l_translation_classes_dependent_data_view(data, lev)
        kwargs["m2l_translation_classes_dependent_data"] = view
        kwargs["translation_classes_level_start"] = cls_start
        kwargs["m2l_translation_classes_lists"] = \
            self.translation_classes_data.from_sep_siblings_translation_classes

    def multipole_to_local(self, actx, level_start_target_box_nrs, target_boxes,
                           src_box_starts, src_box_lists, mpole_exps):
        """M2L of every level.  With translation classes, each run of
        consecutive levels of one order is preprocessed in one launch, the
        levels of an order are contracted in a single MFMA launch
        (:func:`sumpy_amd.e2e.m2l_apply_levels`) and each run is
        postprocessed in one launch (source and target of an M2L share the
        level, so the postprocessing scale ratio is 1 on every level); the
        result equals the reference's per-level loop (``fmm.py:777-920``)."""
        from sumpy_amd.e2e import TILE, m2l_apply_levels
        local_exps = self.local_expansion_zeros(actx)
        starts = _host(level_start_target_box_nrs)
        # translations without preprocessing (the 2D Bessel locals) take the
        # classes' data in the per-level M2L operator itself
        use_pre = (self.supports_translation_classes
                   and self.tree_indep.m2l_translation.use_preprocessing)
        if use_pre:
            pre_mpoles = self.m2l_preproc_mpole_expansion_zeros(actx, mpole_exps)
            work = self.m2l_work_array_zeros(actx, local_exps)
        tr_classes = self.m2l_translation_class_level_start_box_nrs()
        batches = {}   # order -> [(lev, entry, work view)]
        for lev in range(self.tree.nlevels):
            start, stop = starts[lev:lev + 2]
            if start == stop:
                continue
            order = self.level_orders[lev]
            if use_pre:
                if tr_classes[lev] == tr_classes[lev + 1]:
                    continue
                sbase, sview = self.m2l_preproc_mpole_expansions_view(pre_mpoles, lev)
                tbase, tview = self.m2l_work_array_view(work, lev)
            else:
                sbase, sview = self.multipole_expansions_view(mpole_exps, lev)
                tbase, tview = self.local_expansions_view(local_exps, lev)
            m2l = self.tree_indep.m2l(order, order, self.supports_translation_classes)
            kwargs = dict(
                src_expansions=sview, src_base_ibox=sbase, tgt_expansions=tview,
                tgt_base_ibox=tbase, target_boxes=target_boxes[start:stop],
                src_box_starts=src_box_starts[start:stop + 1], src_box_lists=src_box_lists,
                centers=self.tree.box_centers, src_rscale=self.level_to_rscale(lev),
                tgt_rscale=self.level_to_rscale(lev), **self.kernel_extra_kwargs)
            self._add_m2l_precompute_kwargs(kwargs, lev)
            if "m2l_translation_classes_dependent_data" in kwargs and \
                    kwargs["m2l_translation_classes_dependent_data"].numel() == 0:
                continue
            if use_pre:
                batches.setdefault(order, []).append((lev, m2l, kwargs, tview))
            else:
                m2l(actx, **kwargs)
        for order, levs in batches.items():
            runs = _consecutive_runs([lev for lev, _, _, _ in levs])
            mstarts = self.multipole_expansions_level_starts()
            for l0, l1 in runs:
                self.tree_indep.m2l_preprocess_mpole_kernel(order, order)(
                    actx, src_expansions=self._levels_view(mpole_exps, mstarts, l0, l1),
                    preprocessed_src_expansions=self._levels_view(
                        pre_mpoles.flat, pre_mpoles.starts, l0, l1),
                    src_rscale=self.level_to_rscale(l0), **self.kernel_extra_kwargs)
            def entries(tile_width, levs=levs):
                return [op.level_entry(actx, tile_width=tile_width, **kw)
                        for _, op, kw, _ in levs]

            policy, tile_width = None, TILE
            if self.tree_indep.autotune:
                # the split policy and the tile width of the order's batch
                tile_width, policy = self.tree_indep.m2l_tuner.choose(
                    actx, entries, str(self.tree_indep.get_base_kernel()), order,
                    self.tree_indep.m2l_translation.precision)
            m2l_apply_levels(actx, entries(tile_width), policy=policy)
            lstarts = self.local_expansions_level_starts()
            for l0, l1 in runs:
                self.tree_indep.m2l_postprocess_local_kernel(order, order)(
                    actx, tgt_expansions=self._levels_view(local_exps, lstarts, l0, l1),
                    tgt_expansions_before_postprocessing=self._levels_view(
                        work.flat, work.starts, l0, l1),
                    src_rscale=self.level_to_rscale(l0), tgt_rscale=self.level_to_rscale(l0),
                    **self.kernel_extra_kwargs)
        return local_exps

    def eval_multipoles(self, actx, target_boxes_by_source_level, source_boxes_by_level,
                        mpole_exps):
        # host-side knowledge that the result is all zeros (no list-3 work);
        # drive_fmm then skips adding it
        self.last_eval_multipoles_was_empty = all(
            len(t) == 0 for t in target_boxes_by_source_level)
        if self.last_eval_multipoles_was_empty:
            return self._cached_zeros("zero_potentials", lambda: self.output_zeros(actx))
        pot = self.output_zeros(actx)
        kwargs = dict(self.kernel_extra_kwargs)
        kwargs.update(self.box_target_list_kwargs())
        merged = self._list3_all_levels(target_boxes_by_source_level)
        if merged is not None:
            # one launch for every source level (same order on all levels)
            tb, st, li = merged
            m2p = self.tree_indep.m2p(self.level_orders[0])
            res = m2p(actx, src_expansions=self._one_row_per_box(mpole_exps),
                      target_boxes=tb, source_box_starts=st, source_box_lists=li,
                      centers=self.tree.box_centers, result=pot,
                      box_levels=self.tree.box_levels, level_rscales=self._level_rscales(),
                      list34_precision=self.tree_indep.list34_precision, **kwargs)
            for a, b in zip(pot, res):
                assert a is b
            return pot
        for lev, ssn in enumerate(source_boxes_by_level):
            if len(target_boxes_by_source_level[lev]) == 0:
                continue
            m2p = self.tree_indep.m2p(self.level_orders[lev])
            base, view = self.multipole_expansions_view(mpole_exps, lev)
            res = m2p(actx, src_expansions=view, src_base_ibox=base,
                      target_boxes=target_boxes_by_source_level[lev],
                      source_box_starts=ssn.starts, source_box_lists=ssn.lists,
                      centers=self.tree.box_centers, result=pot,
                      rscale=self.level_to_rscale(lev),
                      list34_precision=self.tree_indep.list34_precision, **kwargs)
            for a, b in zip(pot, res):
                assert a is b
        return pot

    def form_locals(self, actx, level_start_target_or_target_parent_box_nrs,
                    target_or_target_parent_boxes, starts, lists, src_weight_vecs):
        lstarts = _host(level_start_target_or_target_parent_box_nrs)
        kwargs = dict(self.extra_kwargs)
        kwargs.update(self.box_source_list_kwargs())
        host_starts = self._host_copy(starts, "from_sep_bigger_starts")
        # host-side knowledge that the result is all zeros (no list-4 work);
        # drive_fmm then skips adding it
        self.last_form_locals_was_empty = host_starts is not None and all(
            lstarts[lev] == lstarts[lev + 1]
            or host_starts[lstarts[lev]] == host_starts[lstarts[lev + 1]]
            for lev in range(self.tree.nlevels))
        if self.last_form_locals_was_empty:
            return self._cached_zeros("zero_locals", lambda: self.local_expansion_zeros(actx))
        local_exps = self.local_expansion_zeros(actx)
        merged = self._list4_all_levels(target_or_target_parent_boxes, starts, host_starts)
        if merged is not None:
            # one launch for every level's target boxes (same order on all levels)
            tb, st = merged
            p2l = self.tree_indep.p2l(self.level_orders[0])
            res = p2l(actx, target_boxes=tb, source_box_starts=st, source_box_lists=lists,
                      centers=self.tree.box_centers, strengths=src_weight_vecs,
                      tgt_expansions=self._one_row_per_box(local_exps), tgt_base_ibox=0,
                      rscale=self.level_to_rscale(0), box_levels=self.tree.box_levels,
                      level_rscales=self._level_rscales(),
                      list34_precision=self.tree_indep.list34_precision, **kwargs)
            return local_exps
        for lev in range(self.tree.nlevels):
            start, stop = lstarts[lev:lev + 2]
            if start == stop:
                continue
            if host_starts is not None and host_starts[start] == host_starts[stop]:
                # no list-4 sources at this level: the zero-initialised
                # expansions are what P2L would write
                continue
            p2l = self.tree_indep.p2l(self.level_orders[lev])
            base, view = self.local_expansions_view(local_exps, lev)
            res = p2l(actx, target_boxes=target_or_target_parent_boxes[start:stop],
                      source_box_starts=starts[start:stop + 1], source_box_lists=lists,
                      centers=self.tree.box_centers, strengths=src_weight_vecs,
                      tgt_expansions=view, tgt_base_ibox=base,
                      rscale=self.level_to_rscale(lev),
                      list34_precision=self.tree_indep.list34_precision, **kwargs)
            assert res is view
        return local_exps

    # {{{ lists 3 and 4 over all levels in one launch

    def _all_levels_ok(self):
        """One launch may span levels when they share the expansion order (one
        row per tree box in the expansion arrays) and the real Laplace kernels
        serve the level (the list 3/4 kernels with per-box rscale)."""
        orders = self.level_orders
        base = self.tree_indep.get_base_kernel()
        return (self.all_levels_lists and len(set(orders)) == 1
                and not base.is_complex_valued and self.tree.dimensions == 3
                and self.tree.nlevels <= 32
                and len(self.tree_indep.local_expansion(orders[0])) == (orders[0] + 1)**2)

    def _l2p_precision(self):
        """FP32 L2P where it is built: potentials (no target derivatives)."""
        from sumpy_amd.kernel import remove_target_transforms
        ti = self.tree_indep
        plain = all(remove_target_transforms(k) == k for k in ti.target_kernels)
        return ti.l2p_precision if plain else "fp64"

    def _level_rscales(self):
        return [self.level_to_rscale(lev) for lev in range(self.tree.nlevels)]

    def _one_row_per_box(self, flat):
        return flat.reshape(self.tree.nboxes, -1)

    def _list3_all_levels(self, target_boxes_by_source_level):
        """(target boxes, starts, source boxes) of list 3 over all source
        levels, a target box's sources level by level (each level's lists in
        their order), or None when the per-level launches must be used."""
        trav = self.traversal
        if not self._all_levels_ok() or "sep_smaller" not in getattr(trav, "host", {}) or \
                target_boxes_by_source_level is not getattr(
                    trav, "target_boxes_sep_smaller_by_source_level", None):
            return None
        memo = self._memo.get("list3_all_levels")
        if memo is None:
            tg, sr = [], []
            for tbs, st, li in trav.host["sep_smaller"]:
                if len(tbs):
                    tg.append(np.repeat(tbs.astype(np.int64), np.diff(st)))
                    sr.append(li[st[0]:st[-1]])
            t = np.concatenate(tg)
            order = np.argsort(t, kind="stable")
            t, src = t[order], np.concatenate(sr)[order]
            boxes, counts = np.unique(t, return_counts=True)
            starts = np.concatenate([[0], np.cumsum(counts)])
            dev = self.tree_indep._setup_actx.from_numpy
            memo = tuple(dev(np.ascontiguousarray(a, np.int32)) for a in (boxes, starts, src))
            self._memo["list3_all_levels"] = memo
        return memo

    def _list4_all_levels(self, totp, starts, host_starts):
        """(target boxes, starts) of the boxes with a non-empty list 4 over all
        levels (their lists stay where they are in the traversal's list), or
        None when the per-level launches must be used."""
        trav = self.traversal
        if not self._all_levels_ok() or host_starts is None or \
                totp is not getattr(trav, "target_or_target_parent_boxes", None):
            return None
        memo = self._memo.get("list4_all_levels")
        if memo is None:
            totp_h = trav.host["target_or_target_parent_boxes"]
            sel = np.nonzero(np.diff(host_starts) > 0)[0]
            # boxes between two selected ones have empty lists, so the
            # selected lists are contiguous: starts[sel] closed by the end of
            # the last one
            st = np.concatenate([host_starts[sel], host_starts[sel[-1] + 1:sel[-1] + 2]])
            dev = self.tree_indep._setup_actx.from_numpy
            memo = (dev(np.ascontiguousarray(totp_h[sel], np.int32)),
                    dev(np.ascontiguousarray(st, np.int32)))
            self._memo["list4_all_levels"] = memo
        return memo

    # }}}

    def refine_locals(self, actx, level_start_target_or_target_parent_box_nrs,
                      target_or_target_parent_boxes, local_exps):
        lstarts = _host(level_start_target_or_target_parent_box_nrs)
        for target_lev in range(1, self.tree.nlevels):
            start, stop = lstarts[target_lev:target_lev + 2]
            if start == stop:
                continue
            source_lev = target_lev - 1
            l2l = self.tree_indep.l2l(self.level_orders[source_lev],
                                      self.level_orders[target_lev])
            sbase, sview = self.local_expansions_view(local_exps, source_lev)
            tbase, tview = self.local_expansions_view(local_exps, target_lev)
            res = l2l(actx, src_expansions=sview, src_base_ibox=sbase,
                      tgt_expansions=tview, tgt_base_ibox=tbase,
                      target_boxes=target_or_target_parent_boxes[start:stop],
                      box_parent_ids=self.tree.box_parent_ids,
                      centers=self.tree.box_centers,
                      src_rscale=self.level_to_rscale(source_lev),
                      tgt_rscale=self.level_to_rscale(target_lev),
                      **self.kernel_extra_kwargs)
            assert res is tview
        return local_exps

    def eval_locals(self, actx, level_start_target_box_nrs, target_boxes, local_exps):
        pot = self.output_zeros(actx)
        starts = _host(level_start_target_box_nrs)
        kwargs = dict(self.kernel_extra_kwargs)
        kwargs.update(self.box_target_list_kwargs())
        if self._all_levels_ok() and target_boxes is getattr(self.traversal, "target_boxes",
                                                              None):
            # one launch for every level's target boxes (same order on all levels)
            l2p = self.tree_indep.l2p(self.level_orders[0])
            res = l2p(actx, src_expansions=self._one_row_per_box(local_exps),
                      target_boxes=target_boxes, centers=self.tree.box_centers, result=pot,
                      box_levels=self.tree.box_levels, level_rscales=self._level_rscales(),
                      l2p_precision=self._l2p_precision(), **kwargs)
            for a, b in zip(pot, res):
                assert a is b
            return pot
        for lev in range(self.tree.nlevels):
            start, stop = starts[lev:lev + 2]
            if start == stop:
                continue
            l2p = self.tree_indep.l2p(self.level_orders[lev])
            base, view = self.local_expansions_view(local_exps, lev)
            res = l2p(actx, src_expansions=view, src_base_ibox=base,
                      target_boxes=target_boxes[start:stop], centers=self.tree.box_centers,
                      result=pot, rscale=self.level_to_rscale(lev), **kwargs)
            for a, b in zip(pot, res):
                assert a is b
        return pot

    def finalize_potentials(self, actx, potentials):
        return potentials

    def _host_copy(self, ary, name):
        """Host copy of a traversal array when *ary* is that array (no
        device synchronisation); None otherwise."""
        trav = self.traversal
        host = getattr(trav, "host", None)
        if host is None or name not in host:
            return None
        try:
            same = ary is getattr(trav, name)
        except AttributeError:
            same = False
        return host[name] if same else None

    # }}}


class _LevelViews(list):
    """Per-level 2D views of one flat allocation (``flat``, with per-level
    entry offsets ``starts``)."""


def _consecutive_runs(levels):
    """Sorted levels -> [(first, last)] of runs of consecutive levels."""
    runs = []
    for lev in sorted(levels):
        if runs and lev == runs[-1][1] + 1:
            runs[-1][1] = lev
        else:
            runs.append([lev, lev])
    return [tuple(r) for r in runs]


def _host(a):
    """Level starts on the host (the reference syncs with ``to_numpy`` here,
    ``fmm.py:614``); host copies kept by the traversal are used when given."""
    if isinstance(a, torch.Tensor):
        return a.cpu().numpy()
    return np.asarray(a)


__all__ = [
    "SumpyExpansionWrangler",
    "SumpyTreeIndependentDataForWrangler",
    "build_csr_level_starts",
]
