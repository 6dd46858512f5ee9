"""QBX layer potential (reference ``sumpy/qbx.py:85-335``).

:class:`LayerPotential` is the dense direct QBX sum over all sources for a
:class:`~sumpy_amd.expansion.local.LineTaylorLocalExpansion`:
``result_k[t] = scale sum_s sigma_s sum_{i<=p} coeff_i(a, b)/i!`` with
``a = c_t - s``, ``b = t - c_t`` (``qbx.py:262-308``), where ``coeff_i`` is the
i-th tau-derivative of the source kernel along ``a + tau b``
(``sumpy/expansion/local.py:132-171``).

Two HIP paths:

* 2D Laplace: the line-Taylor sum in closed complex-analytic form
  (``sumpyhip_qbx2d``; the C5 workload);
* Laplace, Helmholtz and Yukawa in 2D and 3D: power-series arithmetic along
  the line (``sumpyhip_qbx_line``), for the plain kernel, its directional
  or axis source derivative (double layer) and an axis target derivative.
"""

from __future__ import annotations

import numpy as np
import torch

from sumpy_amd import _capi
from sumpy_amd.array_context import obj_array
from sumpy_amd.expansion.local import LineTaylorLocalExpansion
from sumpy_amd.kernel import (
    AxisSourceDerivative,
    AxisTargetDerivative,
    DirectionalSourceDerivative,
    HelmholtzKernel,
    LaplaceKernel,
    YukawaKernel,
)
from sumpy_amd.tools import KernelComputation, as_component_list, as_strength_list

QBX_LINE_MAX_ORDER = 12


def _src_kind(knl):
    """Source-kernel code of the closed-form 2D Laplace path."""
    if isinstance(knl, LaplaceKernel):
        return 0, None
    if isinstance(knl, DirectionalSourceDerivative) and isinstance(knl.inner_kernel, LaplaceKernel):
        return 1, knl.dir_vec_name
    if isinstance(knl, AxisSourceDerivative) and isinstance(knl.inner_kernel, LaplaceKernel):
        return 2 + knl.axis, None
    raise NotImplementedError(f"unsupported QBX source kernel: {knl}")


def _tgt_axis(knl):
    if isinstance(knl, AxisTargetDerivative):
        if knl.inner_kernel != knl.get_base_kernel():
            raise NotImplementedError(f"unsupported QBX target kernel: {knl}")
        return knl.axis
    if knl == knl.get_base_kernel():
        return -1
    raise NotImplementedError(f"unsupported QBX target kernel: {knl}")


def _line_kind(base):
    """Kernel code of ``sumpyhip_qbx_line``."""
    if isinstance(base, LaplaceKernel):
        return 0 if base.dim == 2 else 1
    if isinstance(base, HelmholtzKernel):
        return 2 if base.dim == 2 else 3
    if isinstance(base, YukawaKernel):
        return 4 if base.dim == 2 else 5
    raise NotImplementedError(f"unsupported QBX kernel: {base}")


def _line_src(knl):
    """(source kind, axis, direction-vector name) of ``sumpyhip_qbx_line``."""
    base = knl.get_base_kernel()
    if knl == base:
        return 0, -1, None
    if isinstance(knl, DirectionalSourceDerivative) and knl.inner_kernel == base:
        return 1, -1, knl.dir_vec_name
    if isinstance(knl, AxisSourceDerivative) and knl.inner_kernel == base:
        return 2, knl.axis, None
    raise NotImplementedError(f"unsupported QBX source kernel: {knl}")


class LayerPotentialBase(KernelComputation):
    def __init__(self, expansion, strength_usage=None, value_dtypes=None, name=None,
                 device=None, source_kernels=None, target_kernels=None):
        if source_kernels is None:
            source_kernels = [expansion.kernel]
        if target_kernels is None:
            target_kernels = [expansion.kernel]
        super().__init__(target_kernels, source_kernels, strength_usage, value_dtypes, name)
        self.expansion = expansion
        if not isinstance(expansion, LineTaylorLocalExpansion):
            raise NotImplementedError(
                "LayerPotential is implemented for the LineTaylorLocalExpansion")
        bases = {k.get_base_kernel() for k in (*self.source_kernels, *self.target_kernels)}
        if len(bases) != 1:
            raise ValueError(f"source and target kernels must share one base kernel: {bases}")
        self.base_kernel, = bases
        self.dim = self.base_kernel.dim
        # closed form for 2D Laplace, power series along the line otherwise
        self.closed_form = isinstance(self.base_kernel, LaplaceKernel) and self.dim == 2
        if self.closed_form:
            for k in self.source_kernels:
                _src_kind(k)
        else:
            _line_kind(self.base_kernel)
            for k in self.source_kernels:
                _line_src(k)
            if expansion.order > QBX_LINE_MAX_ORDER:
                raise NotImplementedError(
                    f"line-Taylor order {expansion.order} > {QBX_LINE_MAX_ORDER}")
        for k in self.target_kernels:
            _tgt_axis(k)

    def get_cache_key(self):
        return (type(self).__name__, self.expansion, tuple(self.target_kernels),
                tuple(self.source_kernels), tuple(self.strength_usage),
                tuple(self.value_dtypes))

    def _kernel_param(self, kwargs):
        """(re, im) of the Helmholtz k or the Yukawa lambda (zeros for Laplace)."""
        base = self.base_kernel
        if isinstance(base, HelmholtzKernel):
            k = complex(kwargs[base.helmholtz_k_name])
            if k.imag != 0 and not base.allow_evanescent:
                raise ValueError("complex k needs HelmholtzKernel(allow_evanescent=True)")
            if self.dim == 2 and (k.imag < 0 or (k.imag > 0 and k.real < 0)):
                raise NotImplementedError("the 2D Helmholtz layer potential takes k > 0 "
                                          "or Re k >= 0 < Im k")
            return k.real, k.imag
        if isinstance(base, YukawaKernel):
            return float(kwargs[base.yukawa_lambda_name]), 0.0
        return 0.0, 0.0

    def _coords(self, targets, sources, centers):
        d = self.dim
        tgt, src, ctr = (as_component_list(x, d) for x in (targets, sources, centers))
        return [c + [None]*(3 - d) for c in (tgt, src, ctr)]

    def _finish(self, iknl, res):
        """*res* (complex) in the value dtype of target kernel *iknl*."""
        if np.dtype(self.value_dtypes[iknl]).kind == "c":
            return res
        return res.real.contiguous()


class LayerPotential(LayerPotentialBase):
    """Direct applier for the layer potential (``qbx.py:250-335``)."""

    default_name = "qbx_apply"

    def __call__(self, actx, targets, sources, centers, strengths, expansion_radii,
                 **kwargs):
        dens = []
        for s in as_strength_list(strengths, self.strength_count):
            if s.is_complex():
                raise TypeError("QBX densities must be real")
            dens.append(s.to(torch.float64).contiguous())
        if self.closed_form:
            return self._call_laplace2d(actx, targets, sources, centers, dens, kwargs)
        tgt, src, ctr = self._coords(targets, sources, centers)
        nt, ns = tgt[0].shape[0], src[0].shape[0]
        kind = _line_kind(self.base_kernel)
        kre, kim = self._kernel_param(kwargs)
        scale = complex(self.base_kernel.get_global_scaling_const())
        results = []
        for i, tk in enumerate(self.target_kernels):
            tax = _tgt_axis(tk)
            out = torch.zeros((nt, 2), dtype=torch.float64, device=actx.device)
            for j, sk in enumerate(self.source_kernels):
                skind, sax, dname = _line_src(sk)
                nrm = (as_component_list(kwargs[dname], self.dim) + [None]*(3 - self.dim)
                       if dname else [None]*3)
                _capi.call("sumpyhip_qbx_line", kind, self.expansion.order, skind, sax, tax,
                           kre, kim, nt, *[_capi.ptr(c) for c in tgt],
                           *[_capi.ptr(c) for c in ctr], ns, *[_capi.ptr(c) for c in src],
                           *[_capi.ptr(c) for c in nrm],
                           dens[self.strength_usage[j]].data_ptr(), scale.real, scale.imag,
                           int(j > 0), out.data_ptr(), actx.stream_handle)
            results.append(self._finish(i, torch.view_as_complex(out)))
        return obj_array(results)

    def _call_laplace2d(self, actx, targets, sources, centers, dens, kwargs):
        tgt = as_component_list(targets, 2)
        src = as_component_list(sources, 2)
        ctr = as_component_list(centers, 2)
        nt, ns = tgt[0].shape[0], src[0].shape[0]
        scale = LaplaceKernel(2).get_global_scaling_const()
        results = []
        for i, tk in enumerate(self.target_kernels):
            ax = _tgt_axis(tk)
            out = torch.zeros(nt, dtype=torch.float64, device=actx.device)
            tmp = None
            for j, sk in enumerate(self.source_kernels):
                kind, dname = _src_kind(sk)
                nrm = as_component_list(kwargs[dname], 2) if dname else [None, None]
                dst = out if j == 0 else (tmp if tmp is not None else torch.empty_like(out))
                tmp = dst if j else None
                _capi.call("sumpyhip_qbx2d", self.expansion.order, kind, ax, nt,
                           tgt[0].data_ptr(), tgt[1].data_ptr(), ctr[0].data_ptr(),
                           ctr[1].data_ptr(), ns, src[0].data_ptr(), src[1].data_ptr(),
                           _capi.ptr(nrm[0]), _capi.ptr(nrm[1]),
                           dens[self.strength_usage[j]].data_ptr(), scale, dst.data_ptr(),
                           actx.stream_handle)
                if j:
                    _capi.call("sumpyhip_combine_outputs", 0, nt, 1,
                               _capi.ptr_array([dst.data_ptr()]), _capi.f64_array([1.0]),
                               None, out.data_ptr(), 1, actx.stream_handle)
            if np.dtype(self.value_dtypes[i]).kind == "c":
                out = out.to(torch.complex128)
            results.append(out)
        return obj_array(results)


class _LayerPotentialMatrixBase(LayerPotentialBase):
    """Entries of the layer-potential matrix: each source kernel at unit
    density, summed (reference ``qbx.py:342-539``)."""

    def _entries(self, actx, targets, sources, centers, tgtindices, srcindices, kwargs):
        if self.closed_form:
            return self._entries_laplace2d(actx, targets, sources, centers, tgtindices,
                                           srcindices, kwargs)
        tgt, src, ctr = self._coords(targets, sources, centers)
        nt, ns = tgt[0].shape[0], src[0].shape[0]
        nent, ti, si = self._indices(nt, ns, tgtindices, srcindices)
        kind = _line_kind(self.base_kernel)
        kre, kim = self._kernel_param(kwargs)
        scale = complex(self.base_kernel.get_global_scaling_const())
        results = []
        for i, tk in enumerate(self.target_kernels):
            tax = _tgt_axis(tk)
            out = torch.zeros((nent, 2), dtype=torch.float64, device=actx.device)
            for j, sk in enumerate(self.source_kernels):
                skind, sax, dname = _line_src(sk)
                nrm = (as_component_list(kwargs[dname], self.dim) + [None]*(3 - self.dim)
                       if dname else [None]*3)
                _capi.call("sumpyhip_qbx_line_matrix", kind, self.expansion.order, skind,
                           sax, tax, kre, kim, nt, *[_capi.ptr(c) for c in tgt],
                           *[_capi.ptr(c) for c in ctr], ns, *[_capi.ptr(c) for c in src],
                           *[_capi.ptr(c) for c in nrm], nent, _capi.ptr(ti), _capi.ptr(si),
                           scale.real, scale.imag, int(j > 0), out.data_ptr(),
                           actx.stream_handle)
            results.append(self._finish(i, torch.view_as_complex(out)))
        return results, nt, ns

    @staticmethod
    def _indices(nt, ns, tgtindices, srcindices):
        if tgtindices is None:
            return nt*ns, None, None
        from sumpy_amd.tools import as_i32
        ti, si = as_i32(tgtindices), as_i32(srcindices)
        if ti.shape != si.shape:
            raise ValueError("tgtindices and srcindices must have the same length")
        return int(ti.shape[0]), ti, si

    def _entries_laplace2d(self, actx, targets, sources, centers, tgtindices, srcindices,
                           kwargs):
        tgt = as_component_list(targets, 2)
        src = as_component_list(sources, 2)
        ctr = as_component_list(centers, 2)
        nt, ns = tgt[0].shape[0], src[0].shape[0]
        nent, ti, si = self._indices(nt, ns, tgtindices, srcindices)
        scale = LaplaceKernel(2).get_global_scaling_const()
        results = []
        for i, tk in enumerate(self.target_kernels):
            ax = _tgt_axis(tk)
            out = torch.empty(nent, dtype=torch.float64, device=actx.device)
            for j, sk in enumerate(self.source_kernels):
                kind, dname = _src_kind(sk)
                nrm = as_component_list(kwargs[dname], 2) if dname else [None, None]
                _capi.call("sumpyhip_qbx2d_matrix", self.expansion.order, kind, ax, nt,
                           tgt[0].data_ptr(), tgt[1].data_ptr(), ctr[0].data_ptr(),
                           ctr[1].data_ptr(), ns, src[0].data_ptr(), src[1].data_ptr(),
                           _capi.ptr(nrm[0]), _capi.ptr(nrm[1]), nent, _capi.ptr(ti),
                           _capi.ptr(si), scale, int(j > 0), out.data_ptr(),
                           actx.stream_handle)
            if np.dtype(self.value_dtypes[i]).kind == "c":
                out = out.to(torch.complex128)
            results.append(out)
        return results, nt, ns


class LayerPotentialMatrixGenerator(_LayerPotentialMatrixBase):
    """The ntargets x nsources matrix of each target kernel
    (reference ``qbx.py:342-407``)."""

    default_name = "qbx_matrix"

    def __call__(self, actx, targets, sources, centers, expansion_radii, **kwargs):
        res, nt, ns = self._entries(actx, targets, sources, centers, None, None, kwargs)
        return obj_array([r.view(nt, ns) for r in res])


class LayerPotentialMatrixSubsetGenerator(_LayerPotentialMatrixBase):
    """The entries at the index pairs (tgtindices[i], srcindices[i])
    (reference ``qbx.py:412-539``)."""

    default_name = "qbx_subset"

    def __call__(self, actx, targets, sources, centers, expansion_radii, tgtindices,
                 srcindices, **kwargs):
        res, _, _ = self._entries(actx, targets, sources, centers, tgtindices, srcindices,
                                  kwargs)
        return obj_array(res)


__all__ = ["QBX_LINE_MAX_ORDER", "LayerPotential", "LayerPotentialBase",
           "LayerPotentialMatrixGenerator", "LayerPotentialMatrixSubsetGenerator"]
