"""Particle-to-expansion operators (reference ``sumpy/p2e.py``):
:class:`P2EFromSingleBox` (P2M, ``p2e.py:155-261``) and :class:`P2EFromCSR`
(P2L from list 4, ``p2e.py:268-388``).  Both overwrite the rows of the boxes
they process.
"""

from __future__ import annotations

import torch

from sumpy_amd import _capi, h2d, taylor2d
from sumpy_amd.expansion.local import LocalExpansionBase
from sumpy_amd.kernel import kernel_id, remove_source_transforms, remove_target_transforms
from sumpy_amd.tools import (
    KernelComputation,
    SourceWeights,
    as_component_list,
    as_i32,
    as_strength_list,
    expansion_k,
    helmholtz_k,
    planes,
    set_planes,
)


class P2EBase(KernelComputation):
    def __init__(self, expansion, kernels=None, name=None, strength_usage=None):
        if kernels is None:
            kernels = [remove_target_transforms(expansion.kernel)]
        expansion = expansion.with_kernel(
            remove_source_transforms(remove_target_transforms(expansion.kernel)))
        for knl in kernels:
            if remove_source_transforms(knl) != expansion.kernel:
                raise ValueError(f"kernel {knl} does not match the expansion's kernel")
        super().__init__([], kernels, strength_usage, None, name)
        self.expansion = expansion
        self.dim = expansion.dim

    def get_cache_key(self):
        return (type(self).__name__, self.name, self.expansion,
                tuple(self.source_kernels), tuple(self.strength_usage))

    def _common(self, actx, kwargs):
        sources = as_component_list(kwargs.pop("sources"), self.dim)
        centers = kwargs.pop("centers")
        _capi.require_gpu(centers)
        rscale = float(kwargs.pop("rscale"))
        self._k = expansion_k(self.expansion.kernel, kwargs, self.expansion)
        self._kc = helmholtz_k(self.expansion.kernel, kwargs)
        strengths = as_strength_list(kwargs.pop("strengths"), self.strength_count)
        weights = SourceWeights(actx, self.source_kernels, self.strength_usage, strengths,
                                kwargs, self.dim, self.expansion.is_complex_valued)
        tgt = kwargs.pop("tgt_expansions")
        _capi.require_gpu(tgt)
        if not tgt.is_contiguous():
            raise ValueError("tgt_expansions must be contiguous")
        tables = self.expansion.tables_for(actx.device, self._k, rscale)
        return sources, centers.contiguous(), rscale, weights, tgt, tables


def _ensure_tgt_expansions(op, actx, kwargs):
    """The reference's optional output: without ``tgt_expansions`` a zero
    array of ``nboxes`` rows is made and returned (``sumpy/p2e.py:242-262``)."""
    if kwargs.get("tgt_expansions") is not None:
        return kwargs
    if "nboxes" not in kwargs:
        raise ValueError("either tgt_expansions or nboxes is needed")
    kwargs = dict(kwargs)
    dt = torch.complex128 if op.expansion.is_complex_valued else torch.float64
    kwargs["tgt_expansions"] = torch.zeros((int(kwargs["nboxes"]), len(op.expansion)),
                                           dtype=dt, device=actx.device)
    return kwargs


class P2EFromSingleBox(P2EBase):
    """P2M (reference ``p2e.py:155-261``); with a local expansion, the local
    expansion of each box from its own sources.

    Besides the reference's keywords it takes ``box_levels`` and
    ``level_rscales``: the source boxes may then be of any level (one launch
    for all of them), each with the rscale of its level, and
    ``tgt_expansions`` holds one row per tree box (Laplace kernel);
    ``p2m_precision="fp32"`` then forms the source moments and their box sums
    in FP32 (projected to the stored coefficients in FP64)."""

    default_name = "p2e_from_single_box"

    def __call__(self, actx, **kwargs):
        kwargs = _ensure_tgt_expansions(self, actx, kwargs)
        kwargs.pop("nboxes", None)
        if isinstance(self.expansion, LocalExpansionBase):
            # a local expansion of each box from its own sources: P2L with
            # the box itself as its only list entry
            sb = as_i32(kwargs.pop("source_boxes"))
            kwargs["target_boxes"] = sb
            kwargs["source_box_lists"] = sb
            kwargs["source_box_starts"] = torch.arange(int(sb.shape[0]) + 1, dtype=torch.int32,
                                                       device=sb.device)
            return P2EFromCSR(self.expansion, self.source_kernels,
                              strength_usage=self.strength_usage)(actx, **kwargs)
        if h2d.is_h2d(self.expansion):
            return h2d.p2m(self, actx, kwargs)
        if taylor2d.is_taylor2d(self.expansion):
            return taylor2d.p2m(self, actx, kwargs)
        sources, centers, rscale, weights, tgt, tables = self._common(actx, kwargs)
        source_boxes = as_i32(kwargs.pop("source_boxes"))
        bss = as_i32(kwargs.pop("box_source_starts"))
        bsc = as_i32(kwargs.pop("box_source_counts_nonchild"))
        base = int(kwargs.pop("tgt_base_ibox"))

        levels = kwargs.pop("box_levels", None)
        lrs = kwargs.pop("level_rscales", None)
        prec = {"fp64": 0, "fp32": 1}[kwargs.pop("p2m_precision", "fp64")]
        if prec and levels is None:
            raise ValueError("p2m_precision='fp32' needs box_levels/level_rscales")

        def p2m(wptrs, out):
            if levels is not None:
                # source boxes of several levels in one launch (each with the
                # rscale of its level; one row per tree box)
                _capi.call("sumpyhip_p2m_levels", kernel_id(self.expansion.kernel),
                           tables.desc, int(source_boxes.shape[0]), _capi.ptr(source_boxes),
                           _capi.ptr(bss), _capi.ptr(bsc), _capi.ptr(centers),
                           int(centers.shape[1]), *[s.data_ptr() for s in sources], *wptrs,
                           _capi.ptr(as_i32(levels)), len(lrs), _capi.f64_array(lrs),
                           _capi.ptr(out), prec, actx.stream_handle)
                return
            _capi.call("sumpyhip_p2m", kernel_id(self.expansion.kernel), tables.desc,
                       int(source_boxes.shape[0]), _capi.ptr(source_boxes), _capi.ptr(bss),
                       _capi.ptr(bsc), _capi.ptr(centers), int(centers.shape[1]),
                       *[s.data_ptr() for s in sources], *wptrs, rscale,
                       _capi.ptr(out), base, actx.stream_handle)

        if not tgt.is_complex():
            p2m(weights.ptrs(3), tgt)
            return tgt
        if levels is not None:
            raise NotImplementedError("P2M over several levels: real (Laplace) expansions")
        # complex strengths (Helmholtz): the moments are real-linear in the
        # weights and the projection is real, so the real and imaginary
        # planes are formed one after the other
        w = [weights.w0] + (list(weights.wd) if weights.wd is not None else [])
        re, im = planes(tgt)
        for part, out in ((0, re), (1, im)):
            ws = [torch.view_as_real(x)[..., part].contiguous() for x in w]
            p2m([t.data_ptr() for t in ws] + [None]*(4 - len(ws)), out)
        set_planes(tgt, re, im)
        return tgt


class P2EFromCSR(P2EBase):
    """P2L from list 4 (reference ``p2e.py:268-388``).

    Besides the reference's keywords it takes ``box_levels`` and
    ``level_rscales`` (rscale per level): the target boxes may then be of any
    level (one launch for all of list 4), each with the rscale of its level,
    and ``tgt_expansions`` holds one row per tree box (``rscale`` is then
    only used for the expansion tables); ``list34_precision="fp32"`` then
    forms the terms and their sums over 64 sources in FP32 and adds those
    sums in FP64."""

    default_name = "p2e_from_csr"

    def __call__(self, actx, **kwargs):
        if not isinstance(self.expansion, LocalExpansionBase):
            raise TypeError("P2EFromCSR forms local expansions")
        kwargs = _ensure_tgt_expansions(self, actx, kwargs)
        kwargs.pop("nboxes", None)
        if h2d.is_h2d(self.expansion):
            return h2d.p2l(self, actx, kwargs)
        if taylor2d.is_taylor2d(self.expansion):
            return taylor2d.p2l(self, actx, kwargs)
        prec = {"fp64": 0, "fp32": 1}[kwargs.pop("list34_precision", "fp64")]
        sources, centers, rscale, weights, tgt, tables = self._common(actx, kwargs)
        tb = as_i32(kwargs.pop("target_boxes"))
        starts = as_i32(kwargs.pop("source_box_starts"))
        lists = as_i32(kwargs.pop("source_box_lists"))
        bss = as_i32(kwargs.pop("box_source_starts"))
        bsc = as_i32(kwargs.pop("box_source_counts_nonchild"))
        if tgt.is_complex():
            _capi.call("sumpyhip_p2l_helmholtz", tables.desc, *self._kc,
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                       _capi.ptr(bss), _capi.ptr(bsc), _capi.ptr(centers),
                       int(centers.shape[1]), *[s.data_ptr() for s in sources],
                       *weights.ptrs(3), rscale, _capi.ptr(tgt),
                       int(kwargs.pop("tgt_base_ibox")), prec, actx.stream_handle)
            return tgt
        if "box_levels" in kwargs:
            # target boxes of several levels in one launch (each with the
            # rscale of its level; rows of the whole local array by box)
            levels = as_i32(kwargs.pop("box_levels"))
            lrs = kwargs.pop("level_rscales")
            _capi.call("sumpyhip_p2l_levels", kernel_id(self.expansion.kernel), tables.desc,
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                       _capi.ptr(bss), _capi.ptr(bsc), _capi.ptr(centers),
                       int(centers.shape[1]), *[s.data_ptr() for s in sources],
                       *weights.ptrs(3), _capi.ptr(levels), len(lrs), _capi.f64_array(lrs),
                       _capi.ptr(tgt), prec, actx.stream_handle)
            return tgt
        _capi.call("sumpyhip_p2l", kernel_id(self.expansion.kernel), tables.desc,
                   int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                   _capi.ptr(bss), _capi.ptr(bsc), _capi.ptr(centers),
                   int(centers.shape[1]), *[s.data_ptr() for s in sources],
                   *weights.ptrs(3), rscale, _capi.ptr(tgt),
                   int(kwargs.pop("tgt_base_ibox")), actx.stream_handle)
        return tgt


__all__ = ["P2EFromCSR", "P2EFromSingleBox"]
