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
  :class:`~sumpy_amd.fmm.SumpyExpansionWrangler`).  The strength adjoint is
  the FMM over the same boxes with sources and targets exchanged and the
  conjugate kernel, which is the exact transpose of the discrete forward pass
  (each stage transposes into its exchanged counterpart); source- and
  target-position gradients are the gradients of the adjoint and forward
  fields.  Every expansion is recomputed by the backward passes.

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


__all__ = ["adjoint_wrangler", "fmm_potential", "p2p_potential", "qbx_layer_potential"]


# {{{ FMM (adjoints by the FMM with sources and targets exchanged)

def _check_adjoint_config(wrangler):
    """The adjoint pass computes A^T for the plain single-layer operator:
    one target kernel equal to the base kernel, charge sources with their own
    strength, and self-exclusion only for sources = targets with the identity
    map.  Anything else (derivative kernels, several strengths, another
    target-to-source map) has a different transpose: refuse it."""
    ti = wrangler.tree_indep
    base = ti.get_base_kernel()
    if len(ti.target_kernels) != 1 or ti.target_kernels[0] != base:
        raise NotImplementedError("fmm_potential: one target kernel, the base kernel itself "
                                  f"(got {[str(k) for k in ti.target_kernels]})")
    if ti.source_kernels is not None and list(ti.source_kernels) != [base]:
        raise NotImplementedError("fmm_potential: charge sources only (source_kernels None "
                                  "or [base kernel])")
    if ti.strength_usage is not None and list(ti.strength_usage) != [0]:
        raise NotImplementedError("fmm_potential: one strength (strength_usage None or [0])")
    tree = wrangler.tree
    if ti.exclude_self:
        t2s = wrangler.self_extra_kwargs.get("target_to_source")
        if not tree.sources_are_targets or t2s is None:
            raise NotImplementedError("fmm_potential: exclude_self needs sources = targets")
        ident = torch.arange(tree.ntargets, dtype=t2s.dtype, device=t2s.device)
        if not torch.equal(t2s, ident):
            raise NotImplementedError("fmm_potential: target_to_source must be the identity")


def _conjugate_kernel_kwargs(wrangler):
    """Kernel arguments of the conjugate kernel: Laplace unchanged; Helmholtz
    with a real wavenumber k -> -k (conj e^{ikr}/r = e^{-ikr}/r)."""
    kw = dict(wrangler.kernel_extra_kwargs)
    base = wrangler.tree_indep.get_base_kernel()
    if base.is_complex_valued:
        name = base.get_args()[0]
        k = complex(kw[name])
        if k.imag:
            raise NotImplementedError("the adjoint FMM needs a real wavenumber")
        kw[name] = -k.real
    return kw


def adjoint_wrangler(wrangler, target_kernels=None):
    """A wrangler for the adjoint (transposed, conjugate-kernel) FMM of
    *wrangler*: the same boxes with sources and targets exchanged
    (:func:`sumpy_amd.tree.swapped_traversal`), the same expansion orders,
    precisions and options, the conjugate kernel, and *target_kernels*
    (default: the base kernel; :class:`AxisTargetDerivative` kernels give the
    gradient of the adjoint field, i.e. the source-position gradient).

    Exact transpose: with the roles exchanged every stage is the transpose of
    its forward counterpart up to the parity (-1)^|m| of the coefficients --
    L2P^T = S P2M', M2M^T = S L2L' S, M2L^T = S M2L' S (class matrices
    A_c[t, s] = D(t + s) are symmetric, D_{-d} = S D_d S), M2P^T = S P2L',
    P2P^T = P2P' -- and the signs cancel along every path of the pass, so the
    swapped pass computes A^T g to rounding (for real-valued kernels; with the
    conjugate kernel, A^H g).  Cached on *wrangler* per target-kernel set."""
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.tree import swapped_traversal
    ti = wrangler.tree_indep
    base = ti.get_base_kernel()
    tk = tuple(target_kernels) if target_kernels else (base,)
    key = ("adjoint_wrangler", tk)
    got = wrangler._memo.get(key)
    if got is not None:
        return got
    if tk == tuple(ti.target_kernels) and ti.source_kernels in (None, [base]) and \
            ti.strength_usage in (None, [0]):
        ati = ti
    else:
        ati = ti._memo.get(("adjoint_tree_indep", tk))
        if ati is None:
            ati = SumpyTreeIndependentDataForWrangler(
                ti._setup_actx, ti.multipole_expansion_factory, ti.local_expansion_factory,
                list(tk), exclude_self=ti.exclude_self, use_rscale=ti.use_rscale,
                p2p_precision=ti.p2p_precision, autotune=ti.autotune,
                list34_precision=ti.list34_precision, l2p_precision=ti.l2p_precision,
                p2m_precision=ti.p2m_precision)
            ti._memo[("adjoint_tree_indep", tk)] = ati
    trav = swapped_traversal(ti._setup_actx, wrangler.traversal)
    orders = list(wrangler.level_orders)
    got = SumpyExpansionWrangler(
        ati, trav, wrangler.dtype, fmm_level_to_order=lambda kn, a, t, lev: orders[lev],
        kernel_extra_kwargs=_conjugate_kernel_kwargs(wrangler),
        self_extra_kwargs=wrangler.self_extra_kwargs,
        translation_classes_data=None if trav is not wrangler.traversal
        else wrangler.translation_classes_data,
        preprocessed_mpole_dtype=wrangler.preprocessed_mpole_dtype,
        _disable_translation_classes=not wrangler.supports_translation_classes)
    wrangler._memo[key] = got
    return got


_adjoint_wrangler = adjoint_wrangler


def _gradient_wrangler(wrangler):
    """*wrangler* with the target gradient (AxisTargetDerivative per axis) as
    its target kernels: the target-position gradient."""
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.kernel import AxisTargetDerivative
    key = ("gradient_wrangler",)
    got = wrangler._memo.get(key)
    if got is not None:
        return got
    ti = wrangler.tree_indep
    base = ti.get_base_kernel()
    tk = [AxisTargetDerivative(a, base) for a in range(wrangler.tree.dimensions)]
    gti = SumpyTreeIndependentDataForWrangler(
        ti._setup_actx, ti.multipole_expansion_factory, ti.local_expansion_factory, tk,
        exclude_self=ti.exclude_self, use_rscale=ti.use_rscale,
        p2p_precision=ti.p2p_precision, autotune=ti.autotune,
        list34_precision=ti.list34_precision,
        l2p_precision=ti.l2p_precision, p2m_precision=ti.p2m_precision)
    orders = list(wrangler.level_orders)
    got = SumpyExpansionWrangler(
        gti, wrangler.traversal, wrangler.dtype,
        fmm_level_to_order=lambda kn, a, t, lev: orders[lev],
        kernel_extra_kwargs=wrangler.kernel_extra_kwargs,
        self_extra_kwargs=wrangler.self_extra_kwargs,
        translation_classes_data=wrangler.translation_classes_data,
        preprocessed_mpole_dtype=wrangler.preprocessed_mpole_dtype,
        _disable_translation_classes=not wrangler.supports_translation_classes)
    wrangler._memo[key] = got
    return got


def _re_conj_mul(a, b):
    """Re(conj(a) b), real when both are real."""
    if a.is_complex() or b.is_complex():
        return (torch.conj(a)*b).real if a.is_complex() else a*b.real
    return a*b


class _FMMPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, strength, sources, targets, actx, wrangler, adjoint):
        from sumpy_amd.drive import drive_fmm
        ctx.actx, ctx.wrangler, ctx.adjoint = actx, wrangler, adjoint
        ctx.real_input = not strength.is_complex()
        q = strength.detach().contiguous()
        ctx.save_for_backward(q)
        pot, = drive_fmm(actx, wrangler, [q])
        return pot

    @staticmethod
    def backward(ctx, g):
        from sumpy_amd.drive import drive_fmm
        from sumpy_amd.kernel import AxisTargetDerivative
        q, = ctx.saved_tensors
        wr, actx = ctx.wrangler, ctx.actx
        g = g.contiguous()
        need_q, need_s, need_t = ctx.needs_input_grad[:3]
        dq = ds = dt = None
        if need_q or need_s:
            if need_s:
                base = wr.tree_indep.get_base_kernel()
                dim = wr.tree.dimensions
                awr = adjoint_wrangler(
                    wr, [base] + [AxisTargetDerivative(a, base) for a in range(dim)])
            else:
                awr = ctx.adjoint
            out = drive_fmm(actx, awr, [g])
            if need_q:
                dq = out[0]
                if ctx.real_input and dq.is_complex():
                    dq = dq.real
                dq = dq.contiguous()
            if need_s:
                # dL/ds_j = Re(conj(q_j) grad phi(s_j)), phi = sum_t conj(G)(x - t) g_t
                ds = torch.stack([_re_conj_mul(q, out[1 + a]) for a in range(dim)])
        if need_t:
            gout = drive_fmm(actx, _gradient_wrangler(wr), [q])
            dt = torch.stack([_re_conj_mul(d, g) for d in gout])
        return dq, ds, dt, None, None, None


def fmm_potential(actx, wrangler, strength, adjoint_wrangler=None, sources=None,
                  targets=None):
    """Differentiable FMM potential of *strength* (the caller's order) for
    the wrangler's single target kernel.

    Gradients flow to *strength* and, when given, to *sources* and *targets*:
    ``(dim, n)`` tensors equal to the tree's source / target points in the
    caller's order (the tree itself stays fixed; pass the same tensor as both
    when the sources are the targets to get the total derivative).

    * strengths: the adjoint FMM, i.e. the FMM with sources and targets
      exchanged and the conjugate kernel (:func:`adjoint_wrangler`), the exact
      transpose of the discrete forward pass;
    * source positions: q_j times the gradient of that adjoint field at s_j
      (the same adjoint pass with target-gradient kernels);
    * target positions: g_t times the gradient of the forward field at t
      (the forward pass with target-gradient kernels).

    Every expansion is recomputed by the backward passes; only the strengths
    are kept.  *adjoint_wrangler*: a wrangler for the strength-only adjoint
    pass (default :func:`adjoint_wrangler` of *wrangler*)."""
    from sumpy_amd.tree import user_order_points
    _check_adjoint_config(wrangler)
    tree = wrangler.tree
    for name, x in (("sources", sources), ("targets", targets)):
        if x is not None and not torch.equal(
                x.detach().to(torch.float64), user_order_points(tree, name)):
            raise ValueError(f"fmm_potential: {name} differ from the tree's {name}")
    if adjoint_wrangler is None:
        adjoint_wrangler = _adjoint_wrangler(wrangler)
    return _FMMPotential.apply(strength, sources, targets, actx, wrangler, adjoint_wrangler)


def helmholtz_adjoint_wrangler(wrangler):
    """The strength-adjoint wrangler of a real-wavenumber Helmholtz FMM
    (k -> -k, sources and targets exchanged); :func:`adjoint_wrangler`."""
    return adjoint_wrangler(wrangler)

# }}}
