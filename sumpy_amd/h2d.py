"""Host side of the 2D Hankel/Bessel expansions: Helmholtz (H2D) and
Yukawa (Y2D).

The operators of :mod:`sumpy_amd.p2e`, :mod:`~sumpy_amd.e2e`,
:mod:`~sumpy_amd.e2p` and :mod:`~sumpy_amd.p2p` hand their call here when
their expansion is an H2D / Y2D multipole or local expansion
(:mod:`sumpy_amd.expansion.multipole`, :mod:`~sumpy_amd.expansion.local`),
or their kernel the 2D Helmholtz or Yukawa kernel, with the reference's
keyword contract
(``sumpy/p2e.py:155-388``, ``sumpy/e2e.py:164-1090``, ``sumpy/e2p.py:123-343``,
``sumpy/p2p.py:215-833``).  The HIP kernels are ``csrc/h2d.hip``; the
mathematics is that of ``sumpy/expansion/multipole.py:426-550``,
``sumpy/expansion/local.py:522-704`` and ``sumpy/expansion/m2l.py:1008-1115``
(Yukawa: the Bessel argument scaled by ``i lam``, ``multipole.py:551-567``).

Supported: charges and ``DirectionalSourceDerivative`` / ``AxisSourceDerivative``
dipoles at the sources, potentials and ``AxisTargetDerivative`` gradients at
the targets (with multipliers and sums of those through
:class:`~sumpy_amd.tools.TargetOutputs`), Helmholtz ``k > 0`` or complex ``k`` with ``Im k > 0`` (the
Bessel functions of complex argument then come from the complex recurrences
of ``csrc/bessel.h``), Yukawa ``lam > 0``, orders 0..31, FP64.
"""

from __future__ import annotations

import torch

from sumpy_amd import _capi
from sumpy_amd.tools import (
    SourceWeights,
    TargetOutputs,
    as_component_list,
    as_i32,
    as_strength_list,
    kernel_k,
)

PMAX = 31


def kind_of(kernel) -> int:
    """0: the 2D Helmholtz kernel, 1: the 2D Yukawa kernel, -1: neither."""
    from sumpy_amd.kernel import HelmholtzKernel, YukawaKernel
    base = kernel.get_base_kernel()
    if isinstance(base, HelmholtzKernel) and base.dim == 2:
        return 0
    if isinstance(base, YukawaKernel) and base.dim == 2:
        return 1
    return -1


def is_h2d_kernel(kernel) -> bool:
    """The kernel runs with the 2D Bessel machinery (Helmholtz or Yukawa)."""
    return kind_of(kernel) >= 0


def is_h2d(expansion) -> bool:
    from sumpy_amd.expansion.local import FourierBesselLocalExpansionMixin
    from sumpy_amd.expansion.multipole import HankelBased2DMultipoleExpansion
    return isinstance(expansion, (HankelBased2DMultipoleExpansion,
                                  FourierBesselLocalExpansionMixin))


def wavenumber(kernel, kwargs) -> tuple[float, float]:
    """(Re, Im) of the Helmholtz k, or (lam, 0) for Yukawa, from the call's
    keywords.  A complex Helmholtz k (damped waves, Im k > 0) needs the
    kernel's ``allow_evanescent``, as in the reference's
    ``HelmholtzKernel`` (``sumpy/kernel.py``)."""
    k = kernel_k(kernel, kwargs)
    if kind_of(kernel) == 0 and k.imag != 0:
        if not getattr(kernel.get_base_kernel(), "allow_evanescent", False):
            raise ValueError("a complex Helmholtz k needs HelmholtzKernel(..., "
                             "allow_evanescent=True)")
        if not (k.imag > 0 and k.real >= 0):
            raise NotImplementedError("the 2D Bessel expansions on the GPU take "
                                      "Re k >= 0 and Im k > 0")
        return float(k.real), float(k.imag)
    if k.imag != 0 or not k.real > 0:
        raise NotImplementedError("the 2D Bessel expansions on the GPU need k > 0 or "
                                  "Re k >= 0 < Im k (Helmholtz), lam > 0 (Yukawa)")
    return float(k.real), 0.0


def _order(expansion):
    if not 0 <= expansion.order <= PMAX:
        raise ValueError(f"H2D expansions: order must be in 0..{PMAX}, got {expansion.order}")
    return expansion.order


def _rscale(expansion, r):
    return float(r) if expansion.use_rscale else 1.0


def _complex(t, what):
    _capi.require_gpu(t)
    if t.dtype != torch.complex128 or not t.is_contiguous():
        raise ValueError(f"{what}: a contiguous complex128 tensor is needed")
    return t


def _weights(actx, op, kwargs, strengths):
    """Device pointers (charges, dipole x, dipole y; None without dipoles) and
    the tensors that hold them."""
    w = SourceWeights(actx, op.source_kernels, op.strength_usage,
                      as_strength_list(strengths, op.strength_count), kwargs, 2, True)
    return _weight_ptrs(w)


def _weight_ptrs(w):
    keep = [w.w0.contiguous()] + ([x.contiguous() for x in w.wd] if w.wd is not None else [])
    ptrs = [t.data_ptr() for t in keep] + [None]*(3 - len(keep))
    return ptrs, keep


def _outputs(kernels, result, nt, device):
    """TargetOutputs over potential and (d/dx, d/dy) buffers, and their
    pointers [pot, dx, dy] (None where no kernel needs one)."""
    outs = TargetOutputs(kernels, 2)
    ptrs = outs.buffers(list(result), nt, torch.complex128, device)
    return outs, ptrs[:3]


# {{{ P2M, P2L

def p2m(op, actx, kwargs):
    """``P2EFromSingleBox`` (``sumpy/p2e.py:155-261``): overwrites the rows of
    the source boxes."""
    kwargs = dict(kwargs)
    src = as_component_list(kwargs.pop("sources"), 2)
    centers = kwargs.pop("centers").contiguous()
    tgt = _complex(kwargs.pop("tgt_expansions"), "tgt_expansions")
    k = wavenumber(op.expansion.kernel, kwargs)
    w, _keep = _weights(actx, op, kwargs, kwargs.pop("strengths"))
    sb = as_i32(kwargs.pop("source_boxes"))
    bss = as_i32(kwargs.pop("box_source_starts"))
    bsc = as_i32(kwargs.pop("box_source_counts_nonchild"))
    _capi.call("sumpyhip_h2d_p2m", kind_of(op.expansion.kernel), _order(op.expansion), *k,
               int(sb.shape[0]), _capi.ptr(sb),
               _capi.ptr(bss), _capi.ptr(bsc), _capi.ptr(centers), int(centers.shape[1]),
               src[0].data_ptr(), src[1].data_ptr(), *w,
               _rscale(op.expansion, kwargs.pop("rscale")), _capi.ptr(tgt),
               int(kwargs.pop("tgt_base_ibox")), actx.stream_handle)
    return tgt


def p2l(op, actx, kwargs):
    """``P2EFromCSR`` (``sumpy/p2e.py:268-388``): overwrites the rows of the
    target boxes from their list-4 sources."""
    kwargs = dict(kwargs)
    kwargs.pop("list34_precision", None)
    src = as_component_list(kwargs.pop("sources"), 2)
    centers = kwargs.pop("centers").contiguous()
    tgt = _complex(kwargs.pop("tgt_expansions"), "tgt_expansions")
    k = wavenumber(op.expansion.kernel, kwargs)
    w, _keep = _weights(actx, op, kwargs, kwargs.pop("strengths"))
    tb = as_i32(kwargs.pop("target_boxes"))
    starts = as_i32(kwargs.pop("source_box_starts"))
    lists = as_i32(kwargs.pop("source_box_lists"))
    bss = as_i32(kwargs.pop("box_source_starts"))
    bsc = as_i32(kwargs.pop("box_source_counts_nonchild"))
    _capi.call("sumpyhip_h2d_p2l", kind_of(op.expansion.kernel), _order(op.expansion), *k,
               int(tb.shape[0]), _capi.ptr(tb),
               _capi.ptr(starts), _capi.ptr(lists), _capi.ptr(bss), _capi.ptr(bsc),
               _capi.ptr(centers), int(centers.shape[1]), src[0].data_ptr(),
               src[1].data_ptr(), *w, _rscale(op.expansion, kwargs.pop("rscale")),
               _capi.ptr(tgt), int(kwargs.pop("tgt_base_ibox")), actx.stream_handle)
    return tgt

# }}}


# {{{ M2M, L2L, M2L

def _shift(op, actx, kwargs, local):
    src = _complex(kwargs["src_expansions"], "src_expansions")
    tgt = _complex(kwargs["tgt_expansions"], "tgt_expansions")
    centers = kwargs["centers"].contiguous()
    k = wavenumber(op.tgt_expansion.kernel, kwargs)
    tb = as_i32(kwargs["target_boxes"])
    ps, pt = _order(op.src_expansion), _order(op.tgt_expansion)
    sr = _rscale(op.src_expansion, kwargs["src_rscale"])
    tr = _rscale(op.tgt_expansion, kwargs["tgt_rscale"])
    kind = kind_of(op.tgt_expansion.kernel)
    if local:
        par = as_i32(kwargs["box_parent_ids"])
        _capi.call("sumpyhip_h2d_l2l", kind, ps, pt, *k, int(tb.shape[0]), _capi.ptr(tb),
                   _capi.ptr(par), _capi.ptr(centers), int(centers.shape[1]), sr, tr,
                   _capi.ptr(src), int(kwargs["src_base_ibox"]), _capi.ptr(tgt),
                   int(kwargs["tgt_base_ibox"]), actx.stream_handle)
    else:
        child = as_i32(kwargs["box_child_ids"])
        if child.dim() != 2 or child.shape[0] != 4:
            raise ValueError("box_child_ids must have shape (4, nboxes)")
        _capi.call("sumpyhip_h2d_m2m", kind, ps, pt, *k, int(tb.shape[0]), _capi.ptr(tb),
                   _capi.ptr(child), int(child.shape[1]), _capi.ptr(centers),
                   int(centers.shape[1]), sr, tr, _capi.ptr(src), int(kwargs["src_base_ibox"]),
                   _capi.ptr(tgt), int(kwargs["tgt_base_ibox"]), actx.stream_handle)
    return tgt


def m2m(op, actx, kwargs):
    """``E2EFromChildren`` (``sumpy/e2e.py:866-978``): adds the translated
    children to each target box's row."""
    return _shift(op, actx, kwargs, local=False)


def l2l(op, actx, kwargs):
    """``E2EFromParent`` (``sumpy/e2e.py:985-1090``): adds the translated
    parent local to each target box's row."""
    return _shift(op, actx, kwargs, local=True)


def m2l_data(op, actx, kwargs):
    """``M2LGenerateTranslationClassesDependentData`` (``sumpy/e2e.py:565-669``):
    per class the Hankel row of ``FourierBesselM2LTranslation``
    (``m2l.py:1045-1080``), as (re, im) pairs in FP64."""
    vecs = kwargs["m2l_translation_vectors"].contiguous()
    out = kwargs["m2l_translation_classes_dependent_data"]
    ncls = int(kwargs["ntranslation_classes"])
    ps, pt = _order(op.src_expansion), _order(op.tgt_expansion)
    n = 2*(2*(ps + pt) + 1)
    if tuple(out.shape) != (ncls, n) or out.dtype != torch.float64 or \
            not out.is_contiguous():
        raise ValueError(f"class data must be contiguous FP64 of shape ({ncls}, {n})")
    if int(vecs.shape[0]) != 2:
        raise ValueError("translation vectors must have shape (2, n)")
    _capi.call("sumpyhip_h2d_m2l_data", kind_of(op.tgt_expansion.kernel), ps, pt,
               *wavenumber(op.tgt_expansion.kernel, kwargs),
               ncls, _capi.ptr(vecs), int(vecs.shape[1]),
               int(kwargs["translation_classes_level_start"]), _capi.ptr(out),
               actx.stream_handle)
    return out


def m2l(op, actx, kwargs, with_classes):
    """``E2EFromCSR`` / ``M2LUsingTranslationClassesDependentData``
    (``sumpy/e2e.py:164-558``) with the Fourier-Bessel pre- and
    postprocessing folded in: overwrites each target box's local row."""
    src = _complex(kwargs["src_expansions"], "src_expansions")
    tgt = _complex(kwargs["tgt_expansions"], "tgt_expansions")
    centers = kwargs["centers"].contiguous()
    tb = as_i32(kwargs["target_boxes"])
    starts = as_i32(kwargs["src_box_starts"])
    lists = as_i32(kwargs["src_box_lists"])
    ps, pt = _order(op.src_expansion), _order(op.tgt_expansion)
    classes, data, c0 = None, None, 0
    if with_classes:
        classes = as_i32(kwargs["m2l_translation_classes_lists"])
        data = kwargs["m2l_translation_classes_dependent_data"].contiguous()
        c0 = int(kwargs["translation_classes_level_start"])
        if data.shape[1] != 2*(2*(ps + pt) + 1):
            raise ValueError("class data of the wrong orders")
    _capi.call("sumpyhip_h2d_m2l", kind_of(op.tgt_expansion.kernel), ps, pt,
               *wavenumber(op.tgt_expansion.kernel, kwargs),
               int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
               _capi.ptr(classes), c0, _capi.ptr(data), _capi.ptr(centers),
               int(centers.shape[1]), _rscale(op.src_expansion, kwargs["src_rscale"]),
               _rscale(op.tgt_expansion, kwargs["tgt_rscale"]), _capi.ptr(src),
               int(kwargs["src_base_ibox"]), _capi.ptr(tgt), int(kwargs["tgt_base_ibox"]),
               actx.stream_handle)
    return tgt

# }}}


# {{{ L2P, M2P

def _e2p(op, actx, kwargs, mpole):
    targets = as_component_list(kwargs["targets"], 2)
    centers = kwargs["centers"].contiguous()
    nt = targets[0].shape[0]
    result = kwargs.get("result")
    if result is None:
        from sumpy_amd.array_context import obj_array
        result = obj_array([torch.zeros(nt, dtype=torch.complex128, device=actx.device)
                            for _ in op.kernels])
    outs, out = _outputs(op.kernels, result, nt, actx.device)
    coeffs = _complex(kwargs["src_expansions"], "src_expansions")
    k = wavenumber(op.expansion.kernel, kwargs)
    tb = as_i32(kwargs["target_boxes"])
    bts = as_i32(kwargs["box_target_starts"])
    btc = as_i32(kwargs["box_target_counts_nonchild"])
    p = _order(op.expansion)
    rs = _rscale(op.expansion, kwargs["rscale"])
    kind = kind_of(op.expansion.kernel)
    common = (_capi.ptr(centers), int(centers.shape[1]), targets[0].data_ptr(),
              targets[1].data_ptr(), rs, _capi.ptr(coeffs), int(kwargs["src_base_ibox"]), *out,
              actx.stream_handle)
    if mpole:
        starts = as_i32(kwargs["source_box_starts"])
        lists = as_i32(kwargs["source_box_lists"])
        _capi.call("sumpyhip_h2d_m2p", kind, p, *k, int(tb.shape[0]), _capi.ptr(tb),
                   _capi.ptr(starts), _capi.ptr(lists), _capi.ptr(bts), _capi.ptr(btc), *common)
    else:
        _capi.call("sumpyhip_h2d_l2p", kind, p, *k, int(tb.shape[0]), _capi.ptr(tb),
                   _capi.ptr(bts),
                   _capi.ptr(btc), *common)
    outs.finish(actx, targets, accumulate=mpole)
    return result


def l2p(op, actx, kwargs):
    """``E2PFromSingleBox`` (``sumpy/e2p.py:123-226``): writes the potentials
    of the target boxes' targets."""
    return _e2p(op, actx, kwargs, mpole=False)


def m2p(op, actx, kwargs):
    """``E2PFromCSR`` (``sumpy/e2p.py:233-343``): adds the list-3 multipoles'
    potentials."""
    return _e2p(op, actx, kwargs, mpole=True)

# }}}


# {{{ P2P

def p2p(actx, kernel, kwargs, targets, sources, weights, t2s, target_kernels, result,
        csr=None):
    """``P2P`` (``sumpy/p2p.py:215-285``, dense) or, with *csr* = (target
    boxes, starts, lists, box source starts, counts, box target starts,
    counts), ``P2PFromCSR`` (``sumpy/p2p.py:484-833``) over list 1; writes
    the potentials."""
    k = wavenumber(kernel, kwargs)
    nt = targets[0].shape[0]
    outs, out = _outputs(target_kernels, result, nt, actx.device)
    w, _keep = _weight_ptrs(weights)
    xy = (targets[0].data_ptr(), targets[1].data_ptr())
    sxy = (sources[0].data_ptr(), sources[1].data_ptr())
    if csr is None:
        _capi.call("sumpyhip_h2d_p2p_dense", kind_of(kernel), *k, nt, *xy,
                   int(sources[0].shape[0]), *sxy, *w, _capi.ptr(t2s), *out,
                   actx.stream_handle)
    else:
        ints = [as_i32(a) for a in csr]
        _capi.call("sumpyhip_h2d_p2p_csr", kind_of(kernel), *k, int(ints[0].shape[0]),
                   *[_capi.ptr(a) for a in ints], *xy, *sxy, *w, _capi.ptr(t2s),
                   *out, actx.stream_handle)
    outs.finish(actx, targets)
    return result

# }}}


__all__ = ["is_h2d", "is_h2d_kernel", "kind_of", "l2l", "l2p", "m2l", "m2l_data", "m2m", "m2p",
           "p2l", "p2m", "p2p"]
