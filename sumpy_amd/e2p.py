"""Expansion-to-particle operators (reference ``sumpy/e2p.py``):
:class:`E2PFromSingleBox` (L2P, ``e2p.py:123-226``, overwrites) and
:class:`E2PFromCSR` (M2P from list 3, ``e2p.py:233-343``, accumulates).
"""

from __future__ import annotations

import torch

from sumpy_amd import _capi, h2d, taylor2d
from sumpy_amd.array_context import obj_array
from sumpy_amd.kernel import kernel_id, remove_source_transforms, remove_target_transforms
from sumpy_amd.tools import (
    TargetOutputs,
    as_component_list,
    as_i32,
    expansion_k,
    helmholtz_k,
    planes,
    set_planes,
)


class E2PBase:
    def __init__(self, expansion, kernels, name=None):
        expansion = expansion.with_kernel(remove_source_transforms(expansion.kernel))
        kernels = [remove_source_transforms(k) for k in kernels]
        for knl in kernels:
            if remove_target_transforms(knl) != expansion.kernel:
                raise ValueError(f"kernel {knl} does not match the expansion's kernel")
        self.expansion = expansion
        self.kernels = kernels
        self.name = name or self.default_name
        self.dim = expansion.dim

    @property
    def nresults(self):
        return len(self.kernels)

    def get_cache_key(self):
        return (type(self).__name__, self.expansion, tuple(self.kernels))

    def _setup(self, actx, kwargs):
        targets = as_component_list(kwargs["targets"], self.dim)
        centers = kwargs["centers"].contiguous()
        result = kwargs.get("result")
        nt = targets[0].shape[0]
        if result is None:
            dt = torch.complex128 if self.expansion.is_complex_valued else torch.float64
            result = obj_array([torch.zeros(nt, dtype=dt, device=actx.device)
                                for _ in self.kernels])
        outs = TargetOutputs(self.kernels, self.dim)
        return targets, centers, result, outs


class E2PFromSingleBox(E2PBase):
    """L2P: ``result[t] = scale * sum_mi L_full[mi] (b/rho)^mi/mi!``; a
    multipole expansion is evaluated at its own box's targets (M2P).

    Besides the reference's keywords it takes ``box_levels`` and
    ``level_rscales``: the target boxes may then be of any level (one launch
    for all of them), each with the rscale of its level, and
    ``src_expansions`` holds one row per tree box (Laplace kernel);
    ``l2p_precision="fp32"`` then evaluates the expansion in FP32."""

    default_name = "e2p_from_single_box"

    def __call__(self, actx, **kwargs):
        from sumpy_amd.expansion.local import LocalExpansionBase
        if not isinstance(self.expansion, LocalExpansionBase):
            return self._multipole(actx, kwargs)
        if h2d.is_h2d(self.expansion):
            return h2d.l2p(self, actx, kwargs)
        if taylor2d.is_taylor2d(self.expansion):
            return taylor2d.l2p(self, actx, kwargs)
        targets, centers, result, outs = self._setup(actx, kwargs)
        nt = targets[0].shape[0]
        ptrs = outs.buffers(list(result), nt, result[0].dtype, actx.device)
        rscale = float(kwargs["rscale"] if "rscale" in kwargs else kwargs["level_rscales"][0])
        k = expansion_k(self.expansion.kernel, kwargs, self.expansion)
        tables = self.expansion.tables_for(actx.device, k, rscale)
        tb = as_i32(kwargs["target_boxes"])
        bts = as_i32(kwargs["box_target_starts"])
        btc = as_i32(kwargs["box_target_counts_nonchild"])
        src = kwargs["src_expansions"]

        def l2p(loc, out_ptrs):
            if "box_levels" in kwargs:
                # target boxes of several levels in one launch (each with the
                # rscale of its level; one row per tree box)
                lrs = kwargs["level_rscales"]
                prec = {"fp64": 0, "fp32": 1}[kwargs.get("l2p_precision", "fp64")]
                _capi.call("sumpyhip_l2p_levels", kernel_id(self.expansion.kernel),
                           tables.desc, int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(bts),
                           _capi.ptr(btc), _capi.ptr(centers), int(centers.shape[1]),
                           *[t.data_ptr() for t in targets],
                           _capi.ptr(as_i32(kwargs["box_levels"])), len(lrs),
                           _capi.f64_array(lrs),
                           self.expansion.kernel.get_global_scaling_const(), _capi.ptr(loc),
                           *out_ptrs, prec, actx.stream_handle)
                return
            _capi.call("sumpyhip_l2p", kernel_id(self.expansion.kernel), tables.desc,
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(bts), _capi.ptr(btc),
                       _capi.ptr(centers), int(centers.shape[1]),
                       *[t.data_ptr() for t in targets], rscale,
                       self.expansion.kernel.get_global_scaling_const(),
                       _capi.ptr(loc), int(kwargs["src_base_ibox"]), *out_ptrs,
                       actx.stream_handle)

        if not src.is_complex():
            l2p(src, ptrs)
        elif "box_levels" in kwargs:
            raise NotImplementedError("L2P over several levels: real (Laplace) expansions")
        else:
            # complex locals (Helmholtz): evaluation is real-linear in the
            # coefficients with a real projection; each plane of the locals
            # gives the same plane of the outputs (the targets of other boxes
            # keep their values)
            bufs = outs._bufs + [None]*(4 - len(outs._bufs))
            pl = [planes(b) if b is not None else (None, None) for b in bufs]
            for part, loc in enumerate(planes(src)):
                l2p(loc, [None if b is None else p[part].data_ptr() for b, p in zip(bufs, pl)])
            for b, (re, im) in zip(bufs, pl):
                if b is not None:
                    set_planes(b, re, im)
        # combined target kernels are added to the (zero-initialised) results
        # of the boxes' targets; plain kernels are written in place
        outs.finish(actx, targets, accumulate=not outs.direct)
        return result


    def _multipole(self, actx, kwargs):
        """A multipole expansion evaluated at its own box's targets (the
        reference's E2PFromSingleBox evaluates any expansion): the M2P with
        the box as its only list entry, after clearing those targets (this
        operator overwrites)."""
        kwargs = dict(kwargs)
        tb = as_i32(kwargs["target_boxes"])
        bts = as_i32(kwargs["box_target_starts"])[tb.long()].long()
        btc = as_i32(kwargs["box_target_counts_nonchild"])[tb.long()].long()
        result = kwargs.get("result")
        if result is not None:
            first = torch.repeat_interleave(bts, btc)
            offs = torch.arange(int(btc.sum()), device=tb.device) - torch.repeat_interleave(
                torch.cumsum(btc, 0) - btc, btc)
            idx = first + offs
            for r in result:
                r[idx] = 0
        kwargs["source_box_starts"] = torch.arange(int(tb.shape[0]) + 1, dtype=torch.int32,
                                                   device=tb.device)
        kwargs["source_box_lists"] = tb
        return E2PFromCSR(self.expansion, self.kernels)(actx, **kwargs)


class E2PFromCSR(E2PBase):
    """M2P: ``result[t] += scale * sum_src sum_mi M[mi] rho^|mi| d^mi G(t - c)``.

    Besides the reference's keywords (one source level, ``rscale``,
    ``src_base_ibox``) it takes ``box_levels`` (tree box levels) and
    ``level_rscales`` (rscale per level) instead: the source boxes may then
    be of any level (one launch for all of list 3), each with the rscale of
    its level, and ``src_expansions`` holds one row per tree box;
    ``list34_precision="fp32"`` then evaluates each (target, source box)
    term in FP32 and sums the boxes in FP64."""

    default_name = "e2p_from_csr"

    def __call__(self, actx, **kwargs):
        if h2d.is_h2d(self.expansion):
            return h2d.m2p(self, actx, kwargs)
        if taylor2d.is_taylor2d(self.expansion):
            return taylor2d.m2p(self, actx, kwargs)
        targets, centers, result, outs = self._setup(actx, kwargs)
        nt = targets[0].shape[0]
        ptrs = outs.buffers(list(result), nt, result[0].dtype, actx.device)
        tables = self.expansion.device_tables(actx.device)
        tb = as_i32(kwargs["target_boxes"])
        starts = as_i32(kwargs["source_box_starts"])
        lists = as_i32(kwargs["source_box_lists"])
        bts = as_i32(kwargs["box_target_starts"])
        btc = as_i32(kwargs["box_target_counts_nonchild"])
        if self.expansion.is_complex_valued:
            _capi.call("sumpyhip_m2p_helmholtz", tables.desc,
                       *helmholtz_k(self.expansion.kernel, kwargs),
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                       _capi.ptr(bts), _capi.ptr(btc), _capi.ptr(centers),
                       int(centers.shape[1]), *[t.data_ptr() for t in targets],
                       float(kwargs["rscale"]),
                       self.expansion.kernel.get_global_scaling_const(),
                       _capi.ptr(kwargs["src_expansions"]), int(kwargs["src_base_ibox"]),
                       *ptrs, {"fp64": 0, "fp32": 1}[kwargs.get("list34_precision", "fp64")],
                       actx.stream_handle)
            outs.finish(actx, targets, accumulate=True)
            return result
        if "box_levels" in kwargs:
            # source boxes of several levels in one launch: each box's rscale
            # from its level, rows of the whole multipole array by box number
            levels = as_i32(kwargs["box_levels"])
            rs = _capi.f64_array(kwargs["level_rscales"])
            prec = {"fp64": 0, "fp32": 1}[kwargs.get("list34_precision", "fp64")]
            _capi.call("sumpyhip_m2p_levels", kernel_id(self.expansion.kernel), tables.desc,
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                       _capi.ptr(bts), _capi.ptr(btc), _capi.ptr(centers),
                       int(centers.shape[1]), *[t.data_ptr() for t in targets],
                       _capi.ptr(levels), len(kwargs["level_rscales"]), rs,
                       self.expansion.kernel.get_global_scaling_const(),
                       _capi.ptr(kwargs["src_expansions"]), *ptrs, prec, actx.stream_handle)
        else:
            _capi.call("sumpyhip_m2p", kernel_id(self.expansion.kernel), tables.desc,
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                       _capi.ptr(bts), _capi.ptr(btc), _capi.ptr(centers),
                       int(centers.shape[1]), *[t.data_ptr() for t in targets],
                       float(kwargs["rscale"]),
                       self.expansion.kernel.get_global_scaling_const(),
                       _capi.ptr(kwargs["src_expansions"]), int(kwargs["src_base_ibox"]),
                       *ptrs, actx.stream_handle)
        outs.finish(actx, targets, accumulate=True)
        return result


__all__ = ["E2PFromCSR", "E2PFromSingleBox"]
