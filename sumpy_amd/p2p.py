"""Point-to-point operators (reference ``sumpy/p2p.py``).

:class:`P2P` is the dense direct sum (``p2p.py:215-285``) and
:class:`P2PFromCSR` the FMM near field over per-box source lists
(``p2p.py:484-833``).  Both keep the reference's constructor and keyword
contract and run on the HIP kernels of ``sumpy_amd/csrc/p2p.hip``.

Extra constructor keyword: ``precision`` -- ``"fp64"`` (plain) or ``"fp32"``
(FP32 pair arithmetic with per-target FP64 accumulation).
"""

from __future__ import annotations

import torch

from sumpy_amd import _capi, h2d
from sumpy_amd.array_context import obj_array
from sumpy_amd.kernel import (
    kernel_id,
    remove_source_transforms,
    remove_target_transforms,
)
from sumpy_amd.profiling import kernel_span
from sumpy_amd.tools import (
    KernelComputation,
    SourceWeights,
    TargetOutputs,
    as_component_list,
    as_i32,
    as_strength_list,
    kernel_k,
    new_results,
)

_PRECISIONS = {"fp64": 0, "fp32": 1}


class P2PBase(KernelComputation):
    def __init__(self, target_kernels, exclude_self, strength_usage=None,
                 value_dtypes=None, name=None, source_kernels=None, precision="fp64"):
        if source_kernels is None:
            bases = {remove_target_transforms(k) for k in target_kernels}
            if len(bases) != 1:
                raise ValueError("target kernels must share their source kernel")
            source_kernels = [bases.pop()]
            target_kernels = [remove_source_transforms(k) for k in target_kernels]
        base_src = {k.get_base_kernel() for k in source_kernels}
        base_tgt = {k.get_base_kernel() for k in target_kernels}
        if len(base_src | base_tgt) != 1:
            raise ValueError("source and target kernels must share one base kernel")
        super().__init__(target_kernels, source_kernels, strength_usage, value_dtypes, name)
        self.exclude_self = exclude_self
        self.dim = next(iter(base_src)).dim
        if precision not in _PRECISIONS:
            raise ValueError(f"unknown precision '{precision}'")
        self.precision = precision
        self.base_kernel = next(iter(base_src))

    def _h2d_precision(self):
        if self.precision != "fp64":
            raise NotImplementedError("the 2D Helmholtz P2P runs in FP64 only")

    def get_cache_key(self):
        return (type(self).__name__, tuple(self.target_kernels), self.exclude_self,
                tuple(self.strength_usage), tuple(self.value_dtypes),
                tuple(self.source_kernels), self.precision)

    def _prepare(self, actx, targets, sources, strength, kwargs):
        tgt = as_component_list(targets, self.dim)
        src = as_component_list(sources, self.dim)
        strengths = as_strength_list(strength, self.strength_count)
        is_complex = self.base_kernel.is_complex_valued
        weights = SourceWeights(actx, self.source_kernels, self.strength_usage,
                                strengths, kwargs, self.dim, is_complex)
        t2s = None
        if self.exclude_self:
            t2s = kwargs.get("target_to_source")
            if t2s is None:
                raise ValueError("exclude_self=True needs 'target_to_source'")
            t2s = as_i32(t2s)
        k = kernel_k(self.base_kernel, kwargs)
        return tgt, src, weights, t2s, k


class P2P(P2PBase):
    """Dense direct sum (reference ``p2p.py:215-285``)."""

    default_name = "p2p_apply"

    def __call__(self, actx, targets, sources, strength, **kwargs):
        tgt, src, weights, t2s, k = self._prepare(actx, targets, sources, strength, kwargs)
        nt, ns = tgt[0].shape[0], src[0].shape[0]
        results = new_results(actx, nt, self.value_dtypes)
        if h2d.is_h2d_kernel(self.base_kernel):
            self._h2d_precision()
            return h2d.p2p(actx, self.base_kernel, kwargs, tgt, src, weights, t2s,
                           self.target_kernels, results)
        outs = TargetOutputs(self.target_kernels, self.dim)
        dtype = results[0].dtype
        out_ptrs = outs.buffers(list(results), nt, dtype, actx.device)
        tp = [t.data_ptr() for t in tgt] + [None]*(3 - self.dim)
        sp = [s.data_ptr() for s in src] + [None]*(3 - self.dim)
        scale = self.base_kernel.get_global_scaling_const()
        _capi.call("sumpyhip_p2p_dense", kernel_id(self.base_kernel),
                   _PRECISIONS[self.precision], k.real, k.imag, scale, nt, *tp, ns, *sp,
                   *weights.ptrs(self.dim), _capi.ptr(t2s), *out_ptrs, actx.stream_handle)
        outs.finish(actx, tgt)
        return results


class P2PFromCSR(P2PBase):
    """Near field over per-box CSR source lists (reference ``p2p.py:484-833``).

    Overwrites ``result`` for the targets of ``target_boxes`` and returns the
    same array objects.
    """

    default_name = "p2p_from_csr"

    def __call__(self, actx, targets=None, sources=None, *, max_nsources_in_one_box=None,
                 max_ntargets_in_one_box=None, target_boxes, source_box_starts,
                 source_box_lists, box_source_starts, box_source_counts_nonchild,
                 box_target_starts, box_target_counts_nonchild, strength, result=None,
                 **kwargs):
        tgt, src, weights, t2s, k = self._prepare(actx, targets, sources, strength, kwargs)
        nt = tgt[0].shape[0]
        if result is None:
            result = new_results(actx, nt, self.value_dtypes)
        if h2d.is_h2d_kernel(self.base_kernel):
            self._h2d_precision()
            return h2d.p2p(actx, self.base_kernel, kwargs, tgt, src, weights, t2s,
                           self.target_kernels, result,
                           csr=(target_boxes, source_box_starts, source_box_lists,
                                box_source_starts, box_source_counts_nonchild,
                                box_target_starts, box_target_counts_nonchild))
        outs = TargetOutputs(self.target_kernels, self.dim)
        out_ptrs = outs.buffers(list(result), nt, result[0].dtype, actx.device)
        tp = [t.data_ptr() for t in tgt] + [None]*(3 - self.dim)
        sp = [s.data_ptr() for s in src] + [None]*(3 - self.dim)
        scale = self.base_kernel.get_global_scaling_const()
        keep = [as_i32(a) for a in (target_boxes, source_box_starts, source_box_lists,
                                    box_source_starts, box_source_counts_nonchild,
                                    box_target_starts, box_target_counts_nonchild)]
        ints = [_capi.ptr(a) for a in keep]
        centers = kwargs.get("box_centers")
        nboxes = int(centers.shape[1]) if centers is not None else 0
        with kernel_span(f"p2p_csr_{self.precision}", kwargs.get("_npairs", 0.0)):
            _capi.call("sumpyhip_p2p_csr", kernel_id(self.base_kernel),
                       _PRECISIONS[self.precision], k.real, k.imag, scale,
                       int(target_boxes.shape[0]), *ints, _capi.ptr(centers), nboxes,
                       *tp, *sp, *weights.ptrs(self.dim), _capi.ptr(t2s), *out_ptrs,
                       actx.stream_handle)
        # combined target kernels are added to the (zero-initialised) results
        # of the boxes' targets; plain kernels are written in place
        outs.finish(actx, tgt, accumulate=not outs.direct)
        return result


class _P2PMatrixBase(P2PBase):
    """Matrix entries: each source kernel at unit strength (the reference's
    ``get_strength_or_not`` returns 1), scaled per target kernel; with
    ``exclude_self`` the self pairs are zero."""

    def _entries(self, actx, targets, sources, tgtindices, srcindices, kwargs):
        tgt = as_component_list(targets, self.dim)
        src = as_component_list(sources, self.dim)
        nt, ns = tgt[0].shape[0], src[0].shape[0]
        ones = torch.ones(ns, dtype=torch.float64, device=actx.device)
        is_complex = self.base_kernel.is_complex_valued
        weights = SourceWeights(actx, self.source_kernels, list(range(len(self.source_kernels))),
                                [ones]*len(self.source_kernels), kwargs, self.dim, is_complex)
        t2s = None
        if self.exclude_self:
            t2s = kwargs.get("target_to_source")
            if t2s is None:
                raise ValueError("exclude_self=True needs 'target_to_source'")
            t2s = as_i32(t2s)
        k = kernel_k(self.base_kernel, kwargs)
        if tgtindices is None:
            nent, ti, si = nt*ns, None, None
        else:
            ti, si = as_i32(tgtindices), as_i32(srcindices)
            if ti.shape != si.shape:
                raise ValueError("tgtindices and srcindices must have the same length")
            nent = int(ti.shape[0])
        outs = TargetOutputs(self.target_kernels, self.dim)
        dt = torch.complex128 if is_complex else torch.float64
        raw = [torch.empty(nent, dtype=dt, device=actx.device) if need else None
               for need in [outs.need_pot] + [outs.need_grad]*self.dim]
        ptrs = [None if r is None else torch.view_as_real(r).data_ptr() if is_complex
                else r.data_ptr() for r in raw] + [None]*(3 - self.dim)
        tp = [t.data_ptr() for t in tgt] + [None]*(3 - self.dim)
        sp = [x.data_ptr() for x in src] + [None]*(3 - self.dim)
        _capi.call("sumpyhip_p2p_matrix", kernel_id(self.base_kernel), k.real, k.imag,
                   self.base_kernel.get_global_scaling_const(), nt, *tp, ns, *sp,
                   *weights.ptrs(3), _capi.ptr(t2s), nent, _capi.ptr(ti), _capi.ptr(si),
                   *ptrs, actx.stream_handle)
        # target kernels: linear combinations of the potential and gradient
        # entries, with target-coordinate factors where the kernel has them
        row = None
        results = []
        for ts, vdt in zip(outs.terms, self.value_dtypes):
            acc = None
            for t in ts:
                v = raw[0 if t.deriv < 0 else 1 + t.deriv]*t.coeff
                if t.mult_axis >= 0:
                    if row is None:
                        row = (torch.arange(nt, device=actx.device).repeat_interleave(ns)
                               if ti is None else ti.long())
                    v = v*tgt[t.mult_axis][row]
                acc = v if acc is None else acc + v
            results.append(acc.to(_torch_value_dtype(vdt)))
        return results, nt, ns


def _torch_value_dtype(d):
    from sumpy_amd.tools import _torch_dtype
    return _torch_dtype(d)


class P2PMatrixGenerator(_P2PMatrixBase):
    """The ntargets x nsources interaction matrix of each target kernel
    (reference ``p2p.py:292-345``)."""

    default_name = "p2p_matrix"

    def __call__(self, actx, targets, sources, **kwargs):
        res, nt, ns = self._entries(actx, targets, sources, None, None, kwargs)
        return obj_array([r.view(nt, ns) for r in res])


class P2PMatrixSubsetGenerator(_P2PMatrixBase):
    """The entries at the index pairs (tgtindices[i], srcindices[i])
    (reference ``p2p.py:351-477``)."""

    default_name = "p2p_subset"

    def __call__(self, actx, targets, sources, *, tgtindices, srcindices, **kwargs):
        res, _, _ = self._entries(actx, targets, sources, tgtindices, srcindices, kwargs)
        return obj_array(res)


__all__ = ["P2P", "P2PFromCSR", "P2PMatrixGenerator", "P2PMatrixSubsetGenerator"]
