"""Adjoints of the P2P and QBX direct sums (torch.autograd) against central
finite differences (h = 1e-6, FP64) on random coordinates, and against the
source/target derivative kernels (AxisSourceDerivative / AxisTargetDerivative,
reference kernel.py:2288-2376) evaluated by the forward operators."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

H = 1e-6


def _fd_check(loss_fn, leaves, picks, rtol):
    """Compare autograd gradients of loss_fn() w.r.t. *leaves* with central
    differences at the (leaf index, flat index) pairs *picks*."""
    for x in leaves:
        x.grad = None
    loss = loss_fn()
    loss.backward()
    grads = [x.grad.detach().clone() for x in leaves]
    scale = max(float(g.abs().max()) for g in grads)
    worst = 0.0
    with torch.no_grad():
        for li, fi in picks:
            x = leaves[li].view(-1)
            x0 = float(x[fi])
            x[fi] = x0 + H
            lp = float(loss_fn())
            x[fi] = x0 - H
            lm = float(loss_fn())
            x[fi] = x0
            fd = (lp - lm)/(2*H)
            worst = max(worst, abs(fd - float(grads[li].view(-1)[fi]))/scale)
    assert worst < rtol, worst


def _picks(rng, leaves, n):
    out = []
    for _ in range(n):
        li = int(rng.integers(len(leaves)))
        out.append((li, int(rng.integers(leaves[li].numel()))))
    return out


@pytest.mark.parametrize("case", ["lap3d", "lap3d_self", "lap2d", "helm3d", "helm3d_complex_k"])
def test_p2p_adjoint_fd(actx, case):
    from sumpy_amd.autograd import p2p_potential
    from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel
    rng = np.random.default_rng(11)
    dim = 2 if case == "lap2d" else 3
    ns = 300
    dev = actx.device
    src = torch.tensor(rng.random((dim, ns)), device=dev, requires_grad=True)
    q = torch.tensor(rng.standard_normal(ns), device=dev, requires_grad=True)
    t2s = None
    if case == "lap3d_self":
        tgt = src
        t2s = torch.arange(ns, dtype=torch.int32, device=dev)
    else:
        tgt = torch.tensor(rng.random((dim, 250)) + 0.05, device=dev, requires_grad=True)
    knl = LaplaceKernel(dim) if case.startswith("lap") else HelmholtzKernel(3)
    k = {"helm3d": 5.0, "helm3d_complex_k": 5.0 + 0.7j}.get(case)
    w = torch.tensor(rng.standard_normal(tgt.shape[1]), device=dev)

    def loss():
        y = p2p_potential(tgt, src, q, knl, target_to_source=t2s, k=k)
        return (w*(y.abs()**2 if y.is_complex() else y**2)).sum()

    leaves = [q, src] if case == "lap3d_self" else [q, src, tgt]
    _fd_check(loss, leaves, _picks(rng, leaves, 64), 2e-6)


def test_p2p_adjoint_matches_axis_derivative_kernels(actx):
    """With g = 1: d/dt of y_t is the AxisTargetDerivative potential, and
    d/ds of sum_t y_t is q_s sum_t dG(t - s)/ds, the AxisSourceDerivative
    potential of the single source s summed over the targets."""
    from sumpy_amd import P2P
    from sumpy_amd.autograd import p2p_potential
    from sumpy_amd.kernel import AxisSourceDerivative, AxisTargetDerivative, LaplaceKernel
    rng = np.random.default_rng(5)
    dev = actx.device
    knl = LaplaceKernel(3)
    src = torch.tensor(rng.random((3, 200)), device=dev, requires_grad=True)
    tgt = torch.tensor(rng.random((3, 150)) + 0.1, device=dev, requires_grad=True)
    q = torch.tensor(rng.standard_normal(200), device=dev)
    p2p_potential(tgt, src, q, knl).sum().backward()
    tk = [AxisTargetDerivative(a, knl) for a in range(3)]
    dts = P2P(tk, exclude_self=False)(actx, tgt.detach(), src.detach(), [q])
    for a in range(3):
        assert torch.allclose(tgt.grad[a], dts[a], rtol=1e-12, atol=1e-12)
    for s in (0, 17, 123):
        one = torch.zeros(200, dtype=torch.float64, device=dev)
        one[s] = q[s]
        for a in range(3):
            ys, = P2P([knl], exclude_self=False, source_kernels=[AxisSourceDerivative(a, knl)])(
                actx, tgt.detach(), src.detach(), [one])
            ref = float(ys.sum())
            assert abs(float(src.grad[a, s]) - ref) <= 1e-11*max(1.0, abs(ref))


def _starfish(npanels, order, dev):
    """Gauss-Legendre nodes of r(t) = 1 + 0.25 cos 5t, as a tensor (2, n)."""
    x, _ = np.polynomial.legendre.leggauss(order)
    t = (np.arange(npanels)[:, None] + (x[None, :] + 1)/2).ravel()*2*np.pi/npanels
    r = 1 + 0.25*np.cos(5*t)
    return torch.tensor(np.array([r*np.cos(t), r*np.sin(t)]), device=dev), t


@pytest.mark.parametrize("layer", ["slp", "dlp"])
def test_qbx_adjoint_fd(actx, layer):
    """Gradient of sum_t pot_t^2 w.r.t. the density and the curve nodes
    (sources = targets = nodes, normals from centred differences along the
    curve, centres at half a panel length inside), against central FD."""
    from sumpy_amd.autograd import qbx_layer_potential
    dev = actx.device
    npanels, porder = 32, 16
    X0, t = _starfish(npanels, porder, dev)
    X = X0.clone().requires_grad_(True)
    rng = np.random.default_rng(3)
    sigma = torch.tensor(np.sin(3*t) + 0.01*rng.standard_normal(t.size), device=dev,
                         requires_grad=True)
    wq = 2*np.pi/t.size
    radius = 0.5*2*np.pi/npanels

    def loss():
        tan = torch.roll(X, -1, 1) - torch.roll(X, 1, 1)
        n = torch.stack([tan[1], -tan[0]])/tan.norm(dim=0)
        ctr = X - radius*n
        pot = qbx_layer_potential(X, X, ctr, sigma*wq, 6,
                                  normals=n if layer == "dlp" else None)
        return (pot**2).sum()

    _fd_check(loss, [sigma, X], _picks(rng, [sigma, X], 64), 2e-6)


def test_qbx_adjoint_matches_forward_derivative(actx):
    """The density gradient with g = e_t is row t of the layer potential
    matrix: check it against forward evaluations with unit densities."""
    from sumpy_amd.autograd import qbx_layer_potential
    dev = actx.device
    X, t = _starfish(8, 16, dev)
    n = X/X.norm(dim=0)
    ctr = X - 0.1*n
    sigma = torch.ones(X.shape[1], dtype=torch.float64, device=dev, requires_grad=True)
    pot = qbx_layer_potential(X, X, ctr, sigma, 5, normals=n)
    pot[7].backward()
    for s in (0, 40, 99):
        e = torch.zeros_like(sigma)
        e[s] = 1
        ref = float(qbx_layer_potential(X, X, ctr, e, 5, normals=n)[7])
        assert math.isclose(float(sigma.grad[s]), ref, rel_tol=1e-12, abs_tol=1e-14)


@pytest.mark.parametrize("kernel", ["laplace", "helmholtz"])
def test_fmm_strength_adjoint_fd(actx, kernel):
    """Strength gradient of a whole FMM pass (sources = targets, self
    excluded) by the adjoint pass, against central differences of the FMM
    itself on 64 random strengths.  The FMM is linear in the strengths, so
    the differences are its exact transpose applied to the loss weights; the
    adjoint pass (the FMM of the conjugate kernel) equals it to within the
    FMM's accuracy at order 10 (tolerance 1e-5 relative)."""
    from functools import partial

    from sumpy_amd.autograd import fmm_potential, helmholtz_adjoint_wrangler
    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
    from sumpy_amd.expansion.multipole import LinearPDEConformingVolumeTaylorMultipoleExpansion
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel
    from sumpy_amd.tree import build_tree_and_traversal
    rng = np.random.default_rng(12)
    n = 3000
    src = rng.random((3, n))
    tree, trav = build_tree_and_traversal(actx, src, None, 30)
    knl = LaplaceKernel(3) if kernel == "laplace" else HelmholtzKernel(3)
    cplx = kernel == "helmholtz"
    ti = SumpyTreeIndependentDataForWrangler(
        actx, partial(LinearPDEConformingVolumeTaylorMultipoleExpansion, knl),
        partial(LinearPDEConformingVolumeTaylorLocalExpansion, knl), [knl], exclude_self=True)
    wr = SumpyExpansionWrangler(
        ti, trav, np.complex128 if cplx else np.float64,
        fmm_level_to_order=lambda kn, a, t, lev: 10,
        kernel_extra_kwargs={"k": 3.0} if cplx else None,
        self_extra_kwargs={"target_to_source": actx.from_numpy(np.arange(n, dtype=np.int32))})
    adj = helmholtz_adjoint_wrangler(wr) if cplx else None
    q = rng.standard_normal(n) + (1j*rng.standard_normal(n) if cplx else 0)
    w = rng.standard_normal(n) + (1j*rng.standard_normal(n) if cplx else 0)
    qt = actx.from_numpy(q).requires_grad_(True)
    wt = actx.from_numpy(w)
    pot = fmm_potential(actx, wr, qt, adj)
    loss = (torch.conj(wt)*pot).real.sum() if cplx else (wt*pot).sum()
    loss.backward()
    grad = actx.to_numpy(qt.grad)

    def L(qq):
        p, = drive_fmm(actx, wr, [actx.from_numpy(qq)])
        p = actx.to_numpy(p)
        return float(np.real(np.vdot(w, p))) if cplx else float(w @ p)

    idx = rng.choice(n, 64, replace=False)
    h = 1e-6
    fd_re, fd_im = [], []
    for j in idx:
        e = np.zeros(n, q.dtype)
        e[j] = h
        fd_re.append((L(q + e) - L(q - e))/(2*h))
        if cplx:
            fd_im.append((L(q + 1j*e) - L(q - 1j*e))/(2*h))
    fd = np.array(fd_re) + (1j*np.array(fd_im) if cplx else 0)
    assert np.linalg.norm(grad[idx] - fd)/np.linalg.norm(fd) < 1e-5
