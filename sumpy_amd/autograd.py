"""Adjoints of the direct sums behind :class:`torch.autograd.Function`.

sumpy has no derivatives with respect to strengths or geometry; these are
additions on top of its operators, built from the same HIP kernels:

* :func:`p2p_potential` -- the dense P2P sum ``y_t = scale sum_s q_s G(t - s)``
  (``sumpy/p2p.py:215-285``) for the Laplace 2D/3D and Helmholtz 3D kernels.
  Its backward pass is the transposed sum with the conjugate kernel (for the
  strengths), the gradient of that transposed field at the sources (for the
  source positions) and the gradient of the forward field at the targets
  (for the target positions); nothing from the forward pass is stored but the
  inputs.
* :func:`qbx_layer_potential` -- the 2D Laplace line-Taylor single or double
  layer potential (``sumpy/qbx.py:250-335``); its backward pass is
  ``sumpyhip_qbx2d_adjoint`` (density, sources, normals, targets, centres).

* :func:`fmm_potential` -- a whole FMM pass (``drive_fmm`` over a
  :class:`~sumpy_amd.fmm.SumpyExpansionWrangler`) for sources = targets with
  self-interaction excluded.  The interaction matrix is then symmetric
  (G(t - s) = G(s - t)), so the adjoint for the strengths is one more FMM
  pass with the conjugate kernel (the same wrangler for Laplace; k -> -k for
  a real Helmholtz wavenumber, conj G_k = G_{-k}); every expansion is
  recomputed by that pass, nothing is stored.  It agrees with the exact
  transpose of the discrete FMM operator to within the FMM's accuracy.

Complex outputs follow PyTorch's convention: for a real loss L the incoming
gradient is dL/dRe y + i dL/dIm y, and the gradient of a real input x is
sum_t Re(conj(g_t) dy_t/dx).
"""

from __future__ import annotations

import torch

from sumpy_amd import _capi
from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel, kernel_id

_PREC = {"fp64": 0, "fp32": 1}


def _rows(x, dim):
    x = x.to(torch.float64)
    if x.shape[0] != dim:
        raise ValueError(f"expected coordinates of shape ({dim}, n), got {tuple(x.shape)}")
    return [x[i].contiguous() for i in range(dim)]


def _dense(kid, k, scale, precision, targets, sources, weights, t2s, grad):
    """One sumpyhip_p2p_dense call: potential (and target gradient) of the
    charges *weights* (real, or complex for Helmholtz) at *sources*."""
    dim = targets.shape[0]
    cplx = kid == 13
    tgt, src = _rows(targets, dim), _rows(sources, dim)
    nt = tgt[0].shape[0]
    dev = targets.device
    if cplx:
        w = torch.view_as_real(weights.to(torch.complex128).contiguous()).contiguous()
        pot = torch.zeros(nt, dtype=torch.complex128, device=dev)
        g = [torch.zeros(nt, dtype=torch.complex128, device=dev) for _ in range(dim)] \
            if grad else None
    else:
        w = weights.to(torch.float64).contiguous()
        pot = torch.zeros(nt, dtype=torch.float64, device=dev)
        g = [torch.zeros(nt, dtype=torch.float64, device=dev) for _ in range(dim)] \
            if grad else None
    tp = [t.data_ptr() for t in tgt] + [None]*(3 - dim)
    sp = [s.data_ptr() for s in src] + [None]*(3 - dim)
    gp = ([_capi.ptr(x) for x in g] + [None]*(3 - dim)) if grad else [None]*3
    _capi.call("sumpyhip_p2p_dense", kid, _PREC[precision], float(k.real), float(k.imag),
               scale, nt, *tp, src[0].shape[0], *sp, _capi.ptr(w), None, None, None,
               _capi.ptr(t2s), _capi.ptr(pot), *gp,
               torch.cuda.current_stream(dev).cuda_stream)
    return pot, g


def _inverse_map(t2s, ns):
    """Source -> target map of an exclude-self target_to_source array."""
    if t2s is None:
        return None
    inv = torch.full((ns,), -1, dtype=torch.int32, device=t2s.device)
    ok = t2s >= 0
    inv[t2s[ok].long()] = torch.nonzero(ok).flatten().to(torch.int32)
    return inv


class _P2PPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, strength, sources, targets, kid, k, scale, precision, t2s):
        _capi.require_gpu(strength)
        pot, _ = _dense(kid, k, scale, precision, targets, sources, strength, t2s, False)
        ctx.save_for_backward(strength, sources, targets)
        ctx.args = (kid, k, scale, precision, t2s)
        return pot

    @staticmethod
    def backward(ctx, g):
        strength, sources, targets = ctx.saved_tensors
        kid, k, scale, precision, t2s = ctx.args
        need_q, need_s, need_t = ctx.needs_input_grad[:3]
        dq = ds = dt = None
        cplx_q = strength.is_complex()
        if need_q or need_s:
            # field of the charges g at the targets, conjugate kernel, seen at
            # the sources: sum_t g_t conj(G(t - s)) and its gradient in s
            kc = complex(-k.real, k.imag)  # conj(e^{ikr}) = e^{i(-conj k)r}
            s2t = _inverse_map(t2s, sources.shape[1])
            gin = g if kid == 13 else g.real if g.is_complex() else g
            pot_s, grad_s = _dense(kid, kc, scale, precision, sources, targets, gin, s2t,
                                   need_s)
            if need_q:
                dq = pot_s if cplx_q else pot_s.real if pot_s.is_complex() else pot_s
            if need_s:
                gs = torch.stack(grad_s)
                if gs.is_complex():
                    q = strength.to(torch.complex128)
                    ds = (q[None, :]*gs.conj()).real
                else:
                    ds = strength.to(torch.float64)[None, :]*gs
        if need_t:
            _, grad_t = _dense(kid, k, scale, precision, targets, sources, strength, t2s, True)
            gt = torch.stack(grad_t)
            if gt.is_complex():
                dt = (g.to(torch.complex128).conj()[None, :]*gt).real
            else:
                dt = g.real[None, :]*gt if g.is_complex() else g[None, :]*gt
        return dq, ds, dt, None, None, None, None, None


def p2p_potential(targets, sources, strength, kernel, target_to_source=None, k=None,
                  precision="fp64"):
    """Differentiable dense P2P potential ``scale sum_s q_s G(t - s)``.

    :arg targets, sources: device tensors of shape ``(dim, n)``.
    :arg strength: real (or, for Helmholtz, complex) device tensor ``(nsources,)``.
    :arg kernel: :class:`LaplaceKernel` (2D/3D) or :class:`HelmholtzKernel` (3D).
    :arg target_to_source: optional int32 tensor (exclude-self pairs).
    :arg k: the Helmholtz wave number (complex allowed).
    Gradients flow to *strength*, *sources* and *targets*.
    """
    if isinstance(kernel, HelmholtzKernel):
        if k is None:
            raise ValueError("the Helmholtz kernel needs k")
        k = complex(k)
    elif isinstance(kernel, LaplaceKernel):
        k = 0j
    else:
        raise NotImplementedError(f"p2p_potential: unsupported kernel {kernel}")
    if target_to_source is not None:
        target_to_source = target_to_source.to(torch.int32).contiguous()
    scale = float(kernel.get_global_scaling_const())
    return _P2PPotential.apply(strength, sources, targets, kernel_id(kernel), k, scale,
                               precision, target_to_source)


class _QBXPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, sources, targets, centers, normals, order, scale):
        _capi.require_gpu(density)
        src, tgt, ctr = _rows(sources, 2), _rows(targets, 2), _rows(centers, 2)
        nrm = _rows(normals, 2) if normals is not None else [None, None]
        dens = density.to(torch.float64).contiguous()
        nt = tgt[0].shape[0]
        out = torch.empty(nt, dtype=torch.float64, device=density.device)
        kind = 1 if normals is not None else 0
        _capi.call("sumpyhip_qbx2d", order, kind, -1, nt, _capi.ptr(tgt[0]),
                   _capi.ptr(tgt[1]), _capi.ptr(ctr[0]), _capi.ptr(ctr[1]), src[0].shape[0],
                   _capi.ptr(src[0]), _capi.ptr(src[1]), _capi.ptr(nrm[0]), _capi.ptr(nrm[1]),
                   _capi.ptr(dens), scale, _capi.ptr(out),
                   torch.cuda.current_stream(density.device).cuda_stream)
        ctx.save_for_backward(density, sources, targets, centers,
                              normals if normals is not None else torch.empty(0))
        ctx.args = (order, scale, kind)
        return out

    @staticmethod
    def backward(ctx, g):
        density, sources, targets, centers, normals = ctx.saved_tensors
        order, scale, kind = ctx.args
        need = ctx.needs_input_grad
        src, tgt, ctr = _rows(sources, 2), _rows(targets, 2), _rows(centers, 2)
        nrm = _rows(normals, 2) if kind == 1 else [None, None]
        dev = density.device
        ns, nt = src[0].shape[0], tgt[0].shape[0]

        def pair(n, on):
            return [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(2)] \
                if on else [None, None]
        dd = torch.zeros(ns, dtype=torch.float64, device=dev) if need[0] else None
        ds, dt, dc = pair(ns, need[1]), pair(nt, need[2]), pair(nt, need[3])
        dn = pair(ns, kind == 1 and need[4])
        _capi.call("sumpyhip_qbx2d_adjoint", order, kind, nt, _capi.ptr(tgt[0]),
                   _capi.ptr(tgt[1]), _capi.ptr(ctr[0]), _capi.ptr(ctr[1]), ns,
                   _capi.ptr(src[0]), _capi.ptr(src[1]), _capi.ptr(nrm[0]), _capi.ptr(nrm[1]),
                   _capi.ptr(density.to(torch.float64).contiguous()), scale,
                   _capi.ptr(g.to(torch.float64).contiguous()), _capi.ptr(dd),
                   *[_capi.ptr(x) for x in ds], *[_capi.ptr(x) for x in dn],
                   *[_capi.ptr(x) for x in dt], *[_capi.ptr(x) for x in dc],
                   torch.cuda.current_stream(dev).cuda_stream)

        def st(p):
            return torch.stack(p) if p[0] is not None else None
        return dd, st(ds), st(dt), st(dc), st(dn) if kind == 1 else None, None, None


def qbx_layer_potential(targets, sources, centers, density, order, normals=None):
    """Differentiable 2D Laplace QBX layer potential with a line-Taylor
    expansion of *order* about ``centers`` (one per target): the single layer
    when *normals* is None, else the double layer ``-n . grad_s``
    (``sumpy/kernel.py:2457-2507``), scaled like sumpy's 2D Laplace kernel.
    Gradients flow to *density*, *sources*, *targets*, *centers* and *normals*."""
    scale = float(LaplaceKernel(2).get_global_scaling_const())
    return _QBXPotential.apply(density, sources, targets, centers, normals, int(order), scale)


__all__ = ["p2p_potential", "qbx_layer_potential"]


# {{{ FMM (strength gradient by the adjoint pass)

class _FMMPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, strength, actx, wrangler, adjoint_wrangler):
        from sumpy_amd.drive import drive_fmm
        ctx.actx, ctx.adjoint = actx, adjoint_wrangler
        ctx.real_input = not strength.is_complex()
        pot, = drive_fmm(actx, wrangler, [strength.detach().contiguous()])
        return pot

    @staticmethod
    def backward(ctx, g):
        from sumpy_amd.drive import drive_fmm
        grad, = drive_fmm(ctx.actx, ctx.adjoint, [g.contiguous()])
        if ctx.real_input and grad.is_complex():
            grad = grad.real
        return grad.contiguous(), None, None, None


def fmm_potential(actx, wrangler, strength, adjoint_wrangler=None):
    """Differentiable FMM potential of *strength* (user order) for one target
    kernel, sources = targets, self excluded (the wrangler's
    ``target_to_source``).  *adjoint_wrangler* evaluates the adjoint pass; by
    default the wrangler itself (real kernels).  For the Helmholtz kernel pass
    a wrangler over the same tree with the wavenumber negated
    (:func:`helmholtz_adjoint_wrangler`)."""
    tree = wrangler.tree
    if not tree.sources_are_targets:
        raise NotImplementedError("fmm_potential: sources and targets must coincide")
    if len(wrangler.tree_indep.target_kernels) != 1:
        raise ValueError("fmm_potential: one target kernel")
    if adjoint_wrangler is None:
        if wrangler.tree_indep.get_base_kernel().is_complex_valued:
            raise ValueError("fmm_potential: complex kernels need an adjoint wrangler")
        adjoint_wrangler = wrangler
    return _FMMPotential.apply(strength, actx, wrangler, adjoint_wrangler)


def helmholtz_adjoint_wrangler(wrangler):
    """A wrangler for the adjoint pass of a real-wavenumber Helmholtz FMM:
    same tree, traversal, orders and options, with k -> -k (its translation
    class data are generated for -k)."""
    from sumpy_amd.fmm import SumpyExpansionWrangler
    kw = dict(wrangler.kernel_extra_kwargs)
    name = wrangler.tree_indep.get_base_kernel().get_args()[0]
    k = complex(kw[name])
    if k.imag:
        raise NotImplementedError("the adjoint wrangler needs a real wavenumber")
    kw[name] = -k.real
    orders = list(wrangler.level_orders)
    return SumpyExpansionWrangler(
        wrangler.tree_indep, wrangler.traversal, wrangler.dtype,
        fmm_level_to_order=lambda kn, a, t, lev: orders[lev],
        source_extra_kwargs=wrangler.source_extra_kwargs, kernel_extra_kwargs=kw,
        self_extra_kwargs=wrangler.self_extra_kwargs,
        translation_classes_data=wrangler.translation_classes_data,
        preprocessed_mpole_dtype=wrangler.preprocessed_mpole_dtype)

# }}}
