"""Shared operator plumbing: kernel bookkeeping (reference
``sumpy/tools.py:276-366``, ``KernelComputation``), source-weight assembly and
target-output combination.
"""

from __future__ import annotations

import numpy as np
import torch

from sumpy_amd import _capi
from sumpy_amd.array_context import obj_array
from sumpy_amd.kernel import (
    needs_dipole,
    needs_gradient,
    source_terms,
    target_terms,
    value_dtype,
)


class KernelComputation:
    """Common input processing (reference ``tools.py:276-366``)."""

    def __init__(self, target_kernels, source_kernels, strength_usage=None,
                 value_dtypes=None, name=None):
        if value_dtypes is None:
            value_dtypes = [value_dtype(k) for k in target_kernels]
        if not isinstance(value_dtypes, (list, tuple)):
            value_dtypes = [np.dtype(value_dtypes)]*len(target_kernels)
        if strength_usage is None:
            strength_usage = list(range(len(source_kernels)))
        if len(source_kernels) != len(strength_usage):
            raise ValueError("exprs and strength_usage must have the same length")
        self.source_kernels = tuple(source_kernels)
        self.target_kernels = tuple(target_kernels)
        self.value_dtypes = [np.dtype(v) for v in value_dtypes]
        self.strength_usage = list(strength_usage)
        self.strength_count = max(strength_usage) + 1 if strength_usage else 0
        self.name = name or self.default_name

    @property
    def nresults(self):
        return len(self.target_kernels)


def as_component_list(x, dim):
    """Coordinates given as a (dim, n) tensor or a sequence of dim 1-D tensors
    -> list of contiguous float64 1-D device tensors (no copy when already so)."""
    if isinstance(x, torch.Tensor):
        if x.dim() != 2 or x.shape[0] != dim:
            raise ValueError(f"expected coordinates of shape ({dim}, n), got {tuple(x.shape)}")
        comps = [x[i] for i in range(dim)]
    else:
        comps = list(x)
        if len(comps) != dim:
            raise ValueError(f"expected {dim} coordinate arrays, got {len(comps)}")
    out = []
    for c in comps:
        _capi.require_gpu(c)
        if c.dtype != torch.float64:
            raise TypeError(f"coordinates must be float64, got {c.dtype}")
        out.append(c if c.is_contiguous() else c.contiguous())
    return out


def as_strength_list(strength, count):
    if isinstance(strength, torch.Tensor):
        if strength.dim() == 1:
            strength = [strength]
        else:
            strength = [strength[i] for i in range(strength.shape[0])]
    strength = list(strength)
    if len(strength) < count:
        raise ValueError(f"expected {count} strength arrays, got {len(strength)}")
    for s in strength:
        _capi.require_gpu(s)
    return strength


def _as_weight(t: torch.Tensor, is_complex: bool) -> torch.Tensor:
    dt = torch.complex128 if is_complex else torch.float64
    if t.dtype != dt:
        if not is_complex and t.is_complex():
            raise TypeError("complex strengths for a real kernel")
        t = t.to(dt)
    return t.contiguous()


def _raw(t: torch.Tensor) -> torch.Tensor:
    """Complex tensors as interleaved float64 views."""
    return torch.view_as_real(t) if t.is_complex() else t


class SourceWeights:
    """Charge weights ``w0`` and optional dipole weights ``wd[dim]``, i.e. the
    linear combination of the source kernels (``kernel.py:2288-2507``)."""

    def __init__(self, actx, source_kernels, strength_usage, strengths, kwargs, dim,
                 is_complex):
        terms = [(strength_usage[j], t) for j, k in enumerate(source_kernels)
                 for t in source_terms(k)]
        self.has_dipole = needs_dipole(source_kernels)
        strengths = [_as_weight(s, is_complex) for s in strengths]
        n = strengths[0].shape[0]
        st = actx.stream_handle
        if (len(terms) == 1 and terms[0][1].axis == -1 and terms[0][1].coeff == 1.0):
            self.w0 = strengths[terms[0][0]]
            self.wd = None
            return
        dt = torch.complex128 if is_complex else torch.float64
        dev = strengths[0].device
        self.w0 = torch.zeros(n, dtype=dt, device=dev)
        self.wd = ([torch.zeros(n, dtype=dt, device=dev) for _ in range(dim)]
                   if self.has_dipole else None)
        for iu, t in terms:
            q = strengths[iu]
            if t.dir_vec is not None:
                dvec = kwargs[t.dir_vec]
                dvec = as_component_list(dvec, dim)[t.axis]
            else:
                dvec = None
            target = self.w0 if t.axis < 0 else self.wd[t.axis]
            _capi.call("sumpyhip_axpy_weights", int(is_complex), n, float(t.coeff),
                       _raw(q).data_ptr(), _capi.ptr(dvec), _raw(target).data_ptr(), st)

    def ptrs(self, dim):
        w0 = _raw(self.w0).data_ptr()
        if self.wd is None:
            return [w0, None, None, None]
        wd = [_raw(w).data_ptr() for w in self.wd] + [None]*(3 - dim)
        return [w0, *wd]


class TargetOutputs:
    """Potential/gradient buffers and their combination into the requested
    target kernels (``kernel.py:2330-2588``)."""

    def __init__(self, target_kernels, dim):
        self.dim = dim
        self.terms = [target_terms(k) for k in target_kernels]
        self.need_grad = needs_gradient(target_kernels)
        self.need_pot = any(t.deriv < 0 for ts in self.terms for t in ts)
        # direct: each kernel is exactly one unscaled term without multiplier
        self.direct = (all(len(ts) == 1 and ts[0].coeff == 1.0 and ts[0].mult_axis < 0
                           for ts in self.terms)
                       and len({ts[0].deriv for ts in self.terms}) == len(self.terms))

    def buffers(self, results, n, dtype, device, init=None):
        """Output pointers [pot, g0, g1, g2].  In direct mode they alias the
        result arrays; otherwise temporaries (zero or copies of *init*)."""
        self._results = results
        if self.direct:
            slots = {}
            for r, ts in zip(results, self.terms):
                slots.setdefault(ts[0].deriv, r)
            self._tmp = None
            bufs = [slots.get(-1)] + [slots.get(a) for a in range(self.dim)]
            if self.need_grad:
                for a in range(self.dim):
                    if bufs[1 + a] is None:
                        bufs[1 + a] = torch.zeros(n, dtype=dtype, device=device)
            self._extra_direct = [(r, slots[ts[0].deriv]) for r, ts in zip(results, self.terms)
                                  if slots[ts[0].deriv] is not r]
        else:
            self._tmp = [torch.zeros(n, dtype=dtype, device=device)
                         if (i == 0 and self.need_pot) or (i > 0 and self.need_grad) else None
                         for i in range(1 + self.dim)]
            bufs = self._tmp
        self._bufs = bufs
        ptrs = [None if b is None else _raw(b).data_ptr() for b in bufs]
        return ptrs + [None]*(4 - len(ptrs))

    def finish(self, actx, targets, accumulate=False):
        """Write combined results (non-direct mode); duplicate direct results."""
        if self.direct:
            for r, src in self._extra_direct:
                if accumulate:
                    r += src
                else:
                    r.copy_(src)
            return
        st = actx.stream_handle
        for r, ts in zip(self._results, self.terms):
            ins = [_raw(self._bufs[0 if t.deriv < 0 else 1 + t.deriv]).data_ptr() for t in ts]
            coeffs = [t.coeff for t in ts]
            mults = [None if t.mult_axis < 0 else targets[t.mult_axis].data_ptr() for t in ts]
            _capi.call("sumpyhip_combine_outputs", int(r.is_complex()), r.shape[0], len(ts),
                       _capi.ptr_array(ins), _capi.f64_array(coeffs),
                       _capi.ptr_array(mults), _raw(r).data_ptr(), int(accumulate), st)


def new_results(actx, n, dtypes):
    return obj_array([torch.zeros(n, dtype=_torch_dtype(d), device=actx.device)
                      for d in dtypes])


def _torch_dtype(d):
    from sumpy_amd.array_context import torch_dtype
    return torch_dtype(d)


def as_i32(t: torch.Tensor) -> torch.Tensor:
    """Contiguous int32 device tensor (index arrays are int32 in the
    reference, ``p2p.py:692-695``)."""
    _capi.require_gpu(t)
    if t.dtype != torch.int32:
        t = t.to(torch.int32)
    return t.contiguous()


def kernel_k(kernel, kwargs):
    """Helmholtz parameter from the call kwargs (0 for Laplace)."""
    names = kernel.get_base_kernel().get_args()
    if not names:
        return 0.0
    k = kwargs[names[0]]
    if isinstance(k, torch.Tensor):
        k = k.item()
    from sumpy_amd.kernel import YukawaKernel
    base = kernel.get_base_kernel()
    if isinstance(base, YukawaKernel) and base.dim == 3:
        # e^{-lam r}/r = e^{i k r}/r at k = i lam
        return 1j*complex(k)
    return complex(k)


def expansion_k(kernel, kwargs, expansion=None):
    """Real Helmholtz parameter of an FMM expansion's tables (0 for Laplace).

    The Helmholtz expansions here split complex coefficients into a real and
    an imaginary plane that the real translation kernels process one after
    the other; that needs the PDE projection (the ``(k rho)^2`` terms) to be
    real, hence a real ``k`` for PDE-compressed expansions.  Full-Taylor
    expansions have no projection: with a complex ``k`` (the kernel's
    ``allow_evanescent``) their tables are those of ``k = 0``, and the
    k-dependent operators take :func:`helmholtz_k`."""
    k = kernel_k(kernel, kwargs)
    if k.imag != 0:
        if expansion is None or getattr(expansion, "compressed", True):
            raise NotImplementedError(
                "PDE-compressed Helmholtz expansions on the GPU need a real wavenumber; "
                "use the full-Taylor (VolumeTaylor*) expansions for a complex k")
        _check_evanescent(kernel)
        return 0.0
    return k.real


def _check_evanescent(kernel):
    from sumpy_amd.kernel import YukawaKernel
    if isinstance(kernel.get_base_kernel(), YukawaKernel):
        return
    if not getattr(kernel.get_base_kernel(), "allow_evanescent", False):
        raise ValueError("a complex Helmholtz k needs HelmholtzKernel(..., "
                         "allow_evanescent=True)")


def helmholtz_k(kernel, kwargs):
    """(Re k, Im k) for the k-dependent Helmholtz kernels (class data, M2P,
    P2L, M2L without classes); Im k >= 0."""
    k = kernel_k(kernel, kwargs)
    if k.imag != 0:
        _check_evanescent(kernel)
        if k.imag < 0:
            raise NotImplementedError("complex k needs Im k > 0")
    return float(k.real), float(k.imag)


def planes(z: torch.Tensor):
    """Contiguous real and imaginary planes of a complex tensor."""
    v = torch.view_as_real(z)
    return v[..., 0].contiguous(), v[..., 1].contiguous()


def set_planes(z: torch.Tensor, re: torch.Tensor, im: torch.Tensor):
    """Write real and imaginary planes back into the complex tensor *z*."""
    v = torch.view_as_real(z)
    v[..., 0].copy_(re)
    v[..., 1].copy_(im)
