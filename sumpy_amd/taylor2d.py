"""Host side of the 2D volume-Taylor expansions (full and PDE-compressed) of
the Laplace and Helmholtz kernels.

The operators of :mod:`sumpy_amd.p2e`, :mod:`~sumpy_amd.e2e` and
:mod:`~sumpy_amd.e2p` hand their call here when their expansion is a
(LinearPDEConforming)VolumeTaylor multipole or local expansion of a 2D kernel,
with the reference's keyword contract (``sumpy/p2e.py:155-388``,
``sumpy/e2e.py:164-1090``, ``sumpy/e2p.py:123-343``).  The HIP kernels are
``csrc/taylor2d.hip``; the mathematics is that of
``sumpy/expansion/multipole.py:70-419``, ``sumpy/expansion/local.py:200-510``
and the PDE projection of ``sumpy/expansion/__init__.py:563-886`` in 2D.

Supported: charges and first source derivatives (``DirectionalSourceDerivative``,
``AxisSourceDerivative``) at the sources, potentials and ``AxisTargetDerivative``
gradients (with multipliers, through :class:`~sumpy_amd.tools.TargetOutputs`)
at the targets, Laplace (``log r``, real coefficients) and Helmholtz
(``H0(k r)``, real ``k`` or complex ``k`` with ``Im k > 0``, complex
coefficients), orders 0..12, FP64.
"""

from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np
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

PMAX = 12


class T2DExpn(ctypes.Structure):
    """Mirror of ``struct sumpyhip_t2d_expn``."""
    _fields_ = [("order", ctypes.c_int), ("compressed", ctypes.c_int),
                ("nstored", ctypes.c_int), ("stored_mx", ctypes.c_void_p),
                ("stored_my", ctypes.c_void_p)]


def kind_of(kernel) -> int:
    """0: 2D Laplace, 1: 2D Helmholtz, -1: neither."""
    from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel
    base = kernel.get_base_kernel()
    if base.dim != 2:
        return -1
    if isinstance(base, LaplaceKernel):
        return 0
    if isinstance(base, HelmholtzKernel):
        return 1
    return -1


def is_taylor2d(expansion) -> bool:
    from sumpy_amd.expansion import VolumeTaylorExpansionMixin
    return isinstance(expansion, VolumeTaylorExpansionMixin) and expansion.kernel.dim == 2


def _k(expansion, kwargs) -> tuple[float, float]:
    """(Re k, Im k) of the call (0, 0 for Laplace); a complex Helmholtz k
    needs the kernel's ``allow_evanescent`` and ``Im k > 0``."""
    if kind_of(expansion.kernel) == 0:
        return 0.0, 0.0
    k = kernel_k(expansion.kernel, kwargs)
    if k.imag != 0:
        if not getattr(expansion.kernel.get_base_kernel(), "allow_evanescent", False):
            raise ValueError("a complex Helmholtz k needs HelmholtzKernel(..., "
                             "allow_evanescent=True)")
        if k.imag < 0:
            raise NotImplementedError("the 2D Taylor expansions on the GPU take Im k >= 0")
    elif k.real == 0:
        raise ValueError("Helmholtz k must not be zero")
    return float(k.real), float(k.imag)


@lru_cache(maxsize=None)
def _stored(order: int, compressed: bool, device: str):
    from sumpy_amd.expansion.terms import get_terms_table
    st = get_terms_table(order, 2, compressed).stored
    dev = torch.device(device)
    return (torch.as_tensor(np.ascontiguousarray(st[:, 0]), dtype=torch.int32, device=dev),
            torch.as_tensor(np.ascontiguousarray(st[:, 1]), dtype=torch.int32, device=dev))


def _desc(expansion, device):
    if not 0 <= expansion.order <= PMAX:
        raise ValueError(f"2D Taylor expansions: order must be in 0..{PMAX}, "
                         f"got {expansion.order}")
    mx, my = _stored(expansion.order, bool(expansion.compressed), str(torch.device(device)))
    d = T2DExpn(expansion.order, int(bool(expansion.compressed)), int(mx.shape[0]),
                mx.data_ptr(), my.data_ptr())
    d._keep = (mx, my)
    return d


def _ref(d):
    return ctypes.addressof(d)


def _rows(t, expansion, what):
    _capi.require_gpu(t)
    want = torch.complex128 if kind_of(expansion.kernel) == 1 else torch.float64
    if t.dtype != want or not t.is_contiguous():
        raise ValueError(f"{what}: a contiguous {want} tensor is needed")
    return t


def _rscale(expansion, r):
    return float(r) if expansion.use_rscale else 1.0


def _weights(actx, op, kwargs, strengths, kind):
    """Device pointers (charges, dipole x, dipole y; None without dipoles) and
    the tensors that hold them."""
    w = SourceWeights(actx, op.source_kernels, op.strength_usage,
                      as_strength_list(strengths, op.strength_count), kwargs, 2, kind == 1)
    keep = [w.w0.contiguous()] + ([x.contiguous() for x in w.wd] if w.wd is not None else [])
    return [t.data_ptr() for t in keep] + [None]*(3 - len(keep)), keep


# {{{ P2M, P2L

def p2m(op, actx, kwargs):
    """``P2EFromSingleBox`` (``sumpy/p2e.py:155-261``): overwrites the rows of
    the source boxes."""
    kwargs = dict(kwargs)
    ex = op.expansion
    kind = kind_of(ex.kernel)
    src = as_component_list(kwargs.pop("sources"), 2)
    centers = kwargs.pop("centers").contiguous()
    tgt = _rows(kwargs.pop("tgt_expansions"), ex, "tgt_expansions")
    k = _k(ex, kwargs)
    w, _keep = _weights(actx, op, kwargs, kwargs.pop("strengths"), kind)
    sb = as_i32(kwargs.pop("source_boxes"))
    d = _desc(ex, actx.device)
    _capi.call("sumpyhip_t2d_p2m", kind, _ref(d), *k, int(sb.shape[0]), _capi.ptr(sb),
               _capi.ptr(as_i32(kwargs.pop("box_source_starts"))),
               _capi.ptr(as_i32(kwargs.pop("box_source_counts_nonchild"))),
               _capi.ptr(centers), int(centers.shape[1]), src[0].data_ptr(), src[1].data_ptr(),
               *w, _rscale(ex, kwargs.pop("rscale")), _capi.ptr(tgt),
               int(kwargs.pop("tgt_base_ibox")), actx.stream_handle)
    return tgt


def p2l(op, actx, kwargs):
    """``P2EFromCSR`` (``sumpy/p2e.py:268-388``): overwrites the rows of the
    target boxes from their list-4 sources."""
    kwargs = dict(kwargs)
    kwargs.pop("list34_precision", None)
    ex = op.expansion
    kind = kind_of(ex.kernel)
    src = as_component_list(kwargs.pop("sources"), 2)
    centers = kwargs.pop("centers").contiguous()
    tgt = _rows(kwargs.pop("tgt_expansions"), ex, "tgt_expansions")
    k = _k(ex, kwargs)
    w, _keep = _weights(actx, op, kwargs, kwargs.pop("strengths"), kind)
    tb = as_i32(kwargs.pop("target_boxes"))
    d = _desc(ex, actx.device)
    _capi.call("sumpyhip_t2d_p2l", kind, _ref(d), *k, int(tb.shape[0]), _capi.ptr(tb),
               _capi.ptr(as_i32(kwargs.pop("source_box_starts"))),
               _capi.ptr(as_i32(kwargs.pop("source_box_lists"))),
               _capi.ptr(as_i32(kwargs.pop("box_source_starts"))),
               _capi.ptr(as_i32(kwargs.pop("box_source_counts_nonchild"))),
               _capi.ptr(centers), int(centers.shape[1]), src[0].data_ptr(), src[1].data_ptr(),
               *w, _rscale(ex, kwargs.pop("rscale")), _capi.ptr(tgt),
               int(kwargs.pop("tgt_base_ibox")), actx.stream_handle)
    return tgt

# }}}


# {{{ M2M, L2L, M2L

def _shift(op, actx, kwargs, local):
    se, te = op.src_expansion, op.tgt_expansion
    src = _rows(kwargs["src_expansions"], se, "src_expansions")
    tgt = _rows(kwargs["tgt_expansions"], te, "tgt_expansions")
    centers = kwargs["centers"].contiguous()
    tb = as_i32(kwargs["target_boxes"])
    ds, dt = _desc(se, actx.device), _desc(te, actx.device)
    common = (_capi.ptr(centers), int(centers.shape[1]), _rscale(se, kwargs["src_rscale"]),
              _rscale(te, kwargs["tgt_rscale"]), _capi.ptr(src), int(kwargs["src_base_ibox"]),
              _capi.ptr(tgt), int(kwargs["tgt_base_ibox"]), actx.stream_handle)
    if local:
        par = as_i32(kwargs["box_parent_ids"])
        _capi.call("sumpyhip_t2d_l2l", kind_of(te.kernel), _ref(ds), _ref(dt), *_k(te, kwargs),
                   int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(par), *common)
    else:
        child = as_i32(kwargs["box_child_ids"])
        if child.dim() != 2 or child.shape[0] != 4:
            raise ValueError("box_child_ids must have shape (4, nboxes)")
        _capi.call("sumpyhip_t2d_m2m", kind_of(te.kernel), _ref(ds), _ref(dt), *_k(te, kwargs),
                   int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(child), int(child.shape[1]),
                   *common)
    return tgt


def m2m(op, actx, kwargs):
    """``E2EFromChildren`` (``sumpy/e2e.py:866-978``): adds the translated
    children to each target box's row."""
    return _shift(op, actx, kwargs, local=False)


def l2l(op, actx, kwargs):
    """``E2EFromParent`` (``sumpy/e2e.py:985-1090``): adds the translated
    parent local to each target box's row."""
    return _shift(op, actx, kwargs, local=True)


def class_ndata(tgt_expansion, src_expansion) -> int:
    """FP64 values per translation class: the scaled derivatives of orders up
    to ps + pt (two values each for the complex Helmholtz derivatives)."""
    n = tgt_expansion.order + src_expansion.order
    return (n + 1)*(n + 2)//2*(2 if kind_of(tgt_expansion.kernel) == 1 else 1)


def m2l_data(op, actx, kwargs):
    """``M2LGenerateTranslationClassesDependentData`` (``sumpy/e2e.py:565-669``):
    per class ``rscale^|m| d^m G(v_c)`` for ``|m| <= ps + pt``."""
    vecs = kwargs["m2l_translation_vectors"].contiguous()
    out = kwargs["m2l_translation_classes_dependent_data"]
    ncls = int(kwargs["ntranslation_classes"])
    te, se = op.tgt_expansion, op.src_expansion
    n = class_ndata(te, se)
    if tuple(out.shape) != (ncls, n) or out.dtype != torch.float64 or \
            not out.is_contiguous():
        raise ValueError(f"class data must be contiguous FP64 of shape ({ncls}, {n})")
    if int(vecs.shape[0]) != 2:
        raise ValueError("translation vectors must have shape (2, n)")
    _capi.call("sumpyhip_t2d_m2l_data", kind_of(te.kernel), te.order + se.order,
               *_k(te, kwargs), ncls, _capi.ptr(vecs), int(vecs.shape[1]),
               int(kwargs["translation_classes_level_start"]),
               _rscale(se, kwargs["src_rscale"]), _capi.ptr(out), actx.stream_handle)
    return out


def m2l(op, actx, kwargs, with_classes):
    """``E2EFromCSR`` / ``M2LUsingTranslationClassesDependentData``
    (``sumpy/e2e.py:164-558``): overwrites each target box's local row."""
    se, te = op.src_expansion, op.tgt_expansion
    src = _rows(kwargs["src_expansions"], se, "src_expansions")
    tgt = _rows(kwargs["tgt_expansions"], te, "tgt_expansions")
    centers = kwargs["centers"].contiguous()
    tb = as_i32(kwargs["target_boxes"])
    classes, data, c0 = None, None, 0
    if with_classes:
        classes = as_i32(kwargs["m2l_translation_classes_lists"])
        data = kwargs["m2l_translation_classes_dependent_data"].contiguous()
        c0 = int(kwargs["translation_classes_level_start"])
        if data.shape[1] != class_ndata(te, se):
            raise ValueError("class data of the wrong orders")
    ds, dt = _desc(se, actx.device), _desc(te, actx.device)
    _capi.call("sumpyhip_t2d_m2l", kind_of(te.kernel), _ref(ds), _ref(dt), *_k(te, kwargs),
               int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(as_i32(kwargs["src_box_starts"])),
               _capi.ptr(as_i32(kwargs["src_box_lists"])), _capi.ptr(classes), c0,
               _capi.ptr(data), _capi.ptr(centers), int(centers.shape[1]),
               _rscale(se, kwargs["src_rscale"]), _rscale(te, kwargs["tgt_rscale"]),
               _capi.ptr(src), int(kwargs["src_base_ibox"]), _capi.ptr(tgt),
               int(kwargs["tgt_base_ibox"]), actx.stream_handle)
    return tgt

# }}}


# {{{ L2P, M2P

def _e2p(op, actx, kwargs, mpole):
    ex = op.expansion
    kind = kind_of(ex.kernel)
    targets = as_component_list(kwargs["targets"], 2)
    centers = kwargs["centers"].contiguous()
    nt = targets[0].shape[0]
    dtype = torch.complex128 if kind == 1 else torch.float64
    result = kwargs.get("result")
    if result is None:
        from sumpy_amd.array_context import obj_array
        result = obj_array([torch.zeros(nt, dtype=dtype, device=actx.device)
                            for _ in op.kernels])
    outs = TargetOutputs(op.kernels, 2)
    out = outs.buffers(list(result), nt, dtype, actx.device)[:3]
    coeffs = _rows(kwargs["src_expansions"], ex, "src_expansions")
    tb = as_i32(kwargs["target_boxes"])
    scale = complex(ex.kernel.get_base_kernel().get_global_scaling_const())
    d = _desc(ex, actx.device)
    common = (_capi.ptr(as_i32(kwargs["box_target_starts"])),
              _capi.ptr(as_i32(kwargs["box_target_counts_nonchild"])), _capi.ptr(centers),
              int(centers.shape[1]), targets[0].data_ptr(), targets[1].data_ptr(),
              _rscale(ex, kwargs["rscale"]), _capi.ptr(coeffs), int(kwargs["src_base_ibox"]),
              scale.real, scale.imag, *out, actx.stream_handle)
    if mpole:
        _capi.call("sumpyhip_t2d_m2p", kind, _ref(d), *_k(ex, kwargs), int(tb.shape[0]),
                   _capi.ptr(tb), _capi.ptr(as_i32(kwargs["source_box_starts"])),
                   _capi.ptr(as_i32(kwargs["source_box_lists"])), *common)
    else:
        _capi.call("sumpyhip_t2d_l2p", kind, _ref(d), *_k(ex, kwargs), int(tb.shape[0]),
                   _capi.ptr(tb), *common)
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


__all__ = ["PMAX", "class_ndata", "is_taylor2d", "kind_of", "l2l", "l2p", "m2l", "m2l_data",
           "m2m", "m2p", "p2l", "p2m"]
