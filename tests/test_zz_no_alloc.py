"""The allocation-free M2L contraction (``sumpyhip_m2l_apply_noalloc``,
split partial sums in the library's fixed device scratch) against the
allocating one (``sumpyhip_m2l_apply``, partial sums in buffers the caller
allocates): bitwise equal work rows, on two calls in a row, the second with
a smaller input, in FP64, FP32 and split-BF16; and a call whose partial sums
exceed the scratch fails with an argument error."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _level(actx, n, order, precision, seed):
    """One finest-level M2L contraction entry of an n-point uniform tree."""
    from sumpy_amd import (
        M2LGenerateTranslationClassesDependentData,
        M2LPreprocessMultipole,
        M2LUsingTranslationClassesDependentData,
    )
    from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
    from sumpy_amd.expansion.m2l import VolumeTaylorM2LWithDenseClassMatrices
    from sumpy_amd.expansion.multipole import LinearPDEConformingVolumeTaylorMultipoleExpansion
    from sumpy_amd.kernel import LaplaceKernel
    from sumpy_amd.tree import build_tree_and_traversal
    rng = np.random.default_rng(seed)
    tree, trav = build_tree_and_traversal(actx, rng.random((3, n)), None, 32)
    th, vh = tree.host, trav.host
    knl = LaplaceKernel(3)
    tr = VolumeTaylorM2LWithDenseClassMatrices(precision)
    mp = LinearPDEConformingVolumeTaylorMultipoleExpansion(knl, order)
    lc = LinearPDEConformingVolumeTaylorLocalExpansion(knl, order, m2l_translation_override=tr)
    lbs = th["level_start_box_nrs"]
    lev = tree.nlevels - 1
    nbl = int(lbs[lev + 1] - lbs[lev])
    cls_ls = vh["from_sep_siblings_translation_classes_level_starts"]
    c0, c1 = int(cls_ls[lev]), int(cls_ls[lev + 1])
    vecs = vh["from_sep_siblings_translation_class_to_distance_vector"]
    rs = tree.root_extent*2.0**-lev*2/order
    data = torch.zeros((c1 - c0, tr.translation_classes_dependent_ndata(lc, mp)),
                       dtype=torch.float64, device=actx.device)
    M2LGenerateTranslationClassesDependentData(mp, lc)(
        actx, src_rscale=rs, translation_classes_level_start=c0, ntranslation_classes=c1 - c0,
        m2l_translation_classes_dependent_data=data,
        m2l_translation_vectors=actx.from_numpy(vecs), ntranslation_vectors=vecs.shape[1])
    pdt = torch.float64 if precision == "fp64" else torch.float32
    pre = torch.zeros((nbl, tr.preprocess_multipole_nexprs(lc, mp)), dtype=pdt,
                      device=actx.device)
    M2LPreprocessMultipole(mp, lc)(
        actx, src_expansions=actx.from_numpy(rng.standard_normal((nbl, len(mp)))),
        preprocessed_src_expansions=pre, src_rscale=rs)
    totp = vh["target_or_target_parent_boxes"]
    lst = vh["level_start_target_or_target_parent_box_nrs"]
    s0, s1 = lst[lev], lst[lev + 1]
    op = M2LUsingTranslationClassesDependentData(mp, lc)
    tb = actx.from_numpy(totp[s0:s1])
    starts = actx.from_numpy(vh["from_sep_siblings_starts"][s0:s1 + 1])
    classes = actx.from_numpy(vh["from_sep_siblings_translation_classes"])
    keep = dict(tree=tree, trav=trav, data=data, pre=pre, tb=tb, starts=starts,
                classes=classes)

    def entry(work, tile_width=64):
        return op.level_entry(
            actx, tile_width=tile_width, src_expansions=pre, src_base_ibox=int(lbs[lev]),
            tgt_expansions=work,
            tgt_base_ibox=int(lbs[lev]), target_boxes=tb, src_box_starts=starts,
            src_box_lists=trav.from_sep_siblings_lists, centers=tree.box_centers,
            src_rscale=rs, tgt_rscale=rs, m2l_translation_classes_dependent_data=data,
            translation_classes_level_start=c0, m2l_translation_classes_lists=classes)

    def new_work():
        return torch.zeros((nbl, tr.postprocess_local_nexprs(lc, mp)), dtype=pdt,
                           device=actx.device)
    return entry, new_work, keep


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_m2l_apply_noalloc_matches_allocating(actx, precision):
    from sumpy_amd.e2e import m2l_apply_levels
    for n, seed in ((40000, 0), (6000, 1)):   # the second call on a smaller input
        entry, new_work, keep = _level(actx, n, 6, precision, seed)
        w_alloc, w_noalloc = new_work(), new_work()
        # 8 items per CU with at least 6 slots per split: several splits per tile
        ent = entry(w_alloc)
        m2l_apply_levels(actx, [ent], policy=(8, 6))
        m2l_apply_levels(actx, [entry(w_noalloc)], policy=(8, 6), noalloc=True)
        torch.cuda.synchronize()
        assert torch.count_nonzero(w_alloc).item() > 0
        assert torch.equal(w_alloc, w_noalloc), (precision, n)


def test_m2l_apply_noalloc_rejects_oversized_partial_sums(actx):
    """More split partial sums than the fixed scratch holds: an argument
    error (ValueError), no launch."""
    from sumpy_amd.e2e import m2l_apply_levels
    entry, new_work, keep = _level(actx, 40000, 6, "fp64", 0)
    work = new_work()
    ent = entry(work)
    sched = dict(ent["sched"], nsplit=4096)   # far beyond the scratch
    with pytest.raises(ValueError, match="fixed scratch"):
        m2l_apply_levels(actx, [dict(ent, sched=sched)], noalloc=True)
