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
    adjoint pass (the FMM of the conjugate kernel over the exchanged roles) is
    that transpose to rounding; the differences themselves carry the
    rounding of L (~1e-16 |L| / h), about 1e-8 relative here (tolerance
    1e-7)."""
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
    assert np.linalg.norm(grad[idx] - fd)/np.linalg.norm(fd) < 1e-7


def _fmm_setup(actx, kernel, src, tgt, order, source_kernels=None, target_kernels=None,
               tree_trav=None):
    from functools import partial

    from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
    from sumpy_amd.expansion.multipole import LinearPDEConformingVolumeTaylorMultipoleExpansion
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel
    from sumpy_amd.tree import build_tree_and_traversal
    knl = LaplaceKernel(3) if kernel == "laplace" else HelmholtzKernel(3)
    cplx = kernel == "helmholtz"
    tree, trav = tree_trav or build_tree_and_traversal(actx, src, tgt, 24)
    same = tree.sources_are_targets   # then the self pair is excluded
    ti = SumpyTreeIndependentDataForWrangler(
        actx, partial(LinearPDEConformingVolumeTaylorMultipoleExpansion, knl),
        partial(LinearPDEConformingVolumeTaylorLocalExpansion, knl),
        target_kernels or [knl], source_kernels=source_kernels, exclude_self=same)
    wr = SumpyExpansionWrangler(
        ti, trav, np.complex128 if cplx else np.float64,
        fmm_level_to_order=lambda kn, a, t, lev: order,
        kernel_extra_kwargs={"k": 2.5} if cplx else None,
        self_extra_kwargs={"target_to_source": actx.from_numpy(
            np.arange(tree.ntargets, dtype=np.int32))} if same else None)
    return knl, tree, trav, wr


def _separate(seed, ns=4000, nt=1500):
    """Clustered sources (lists 3 and 4 populated) and separate uniform
    targets."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.25, 0.75, (4, 3))
    src = np.clip(c[rng.integers(0, 4, ns)].T + 0.05*rng.standard_normal((3, ns)), 0, 1)
    return rng, src, rng.random((3, nt))


@pytest.mark.parametrize("kernel", ["laplace", "helmholtz"])
def test_fmm_adjoint_is_exact_transpose_separate_targets(actx, kernel):
    """Sources != targets: the adjoint pass (roles exchanged, conjugate
    kernel) satisfies <g, A q> = <A^H g, q> to rounding, on a tree whose
    lists 3 and 4 are populated; and the strength gradient matches central
    differences of the FMM on 64 strengths."""
    from sumpy_amd.autograd import fmm_potential
    from sumpy_amd.drive import drive_fmm
    rng, src, tgt = _separate(21)
    knl, tree, trav, wr = _fmm_setup(actx, kernel, src, tgt, 8)
    assert sum(len(x[2]) for x in trav.host["sep_smaller"]) > 0
    assert len(trav.host["from_sep_bigger_lists"]) > 0
    cplx = kernel == "helmholtz"
    ns, nt = src.shape[1], tgt.shape[1]
    q = rng.standard_normal(ns) + (1j*rng.standard_normal(ns) if cplx else 0)
    g = rng.standard_normal(nt) + (1j*rng.standard_normal(nt) if cplx else 0)
    qt = actx.from_numpy(q).requires_grad_(True)
    pot = fmm_potential(actx, wr, qt)
    gt = actx.from_numpy(g)
    loss = (torch.conj(gt)*pot).real.sum()
    loss.backward()
    aq = actx.to_numpy(pot.detach())
    ahg = actx.to_numpy(qt.grad)          # A^H g (real part only for real q)
    lhs, rhs = np.vdot(g, aq), np.vdot(ahg, q)
    assert abs(lhs - rhs) <= 1e-12*np.linalg.norm(g)*np.linalg.norm(aq), (lhs, rhs)

    def L(qq):
        p, = drive_fmm(actx, wr, [actx.from_numpy(qq)])
        return float(np.real(np.vdot(g, actx.to_numpy(p))))
    idx = rng.choice(ns, 64, replace=False)
    h = 1e-6
    fd = []
    for j in idx:
        e = np.zeros(ns, q.dtype)
        e[j] = h
        v = (L(q + e) - L(q - e))/(2*h)
        if cplx:
            v = v + 1j*(L(q + 1j*e) - L(q - 1j*e))/(2*h)
        fd.append(v)
    fd = np.array(fd)
    # the differences carry the rounding of L over 2h (~1e-8 relative)
    assert np.linalg.norm(ahg[idx] - fd)/np.linalg.norm(fd) < 1e-7


def _tree_index(tree, which):
    """tree position of each point in the caller's order"""
    if which == "sources":
        inv = np.empty_like(tree.host["user_source_ids"])
        inv[tree.host["user_source_ids"]] = np.arange(len(inv))
        return inv
    return tree.host["sorted_target_ids"]


@pytest.mark.parametrize("kernel,same", [("laplace", False), ("laplace", True),
                                         ("helmholtz", False)])
def test_fmm_position_gradients_fd(actx, kernel, same):
    """Source- and target-position gradients of sum_t Re(conj(w_t) pot_t)
    through the whole FMM against central differences (h = 1e-6) of the FMM
    with the points moved in place (the tree fixed), on 64 coordinates;
    sources = targets passes one tensor for both (total derivative)."""
    from sumpy_amd.autograd import fmm_potential
    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.tree import user_order_points
    rng, src, tgt = _separate(22)
    knl, tree, trav, wr = _fmm_setup(actx, kernel, src, None if same else tgt, 8)
    cplx = kernel == "helmholtz"
    ns, nt = tree.nsources, tree.ntargets
    q = rng.standard_normal(ns) + (1j*rng.standard_normal(ns) if cplx else 0)
    w = rng.standard_normal(nt) + (1j*rng.standard_normal(nt) if cplx else 0)
    S = user_order_points(tree, "sources").clone().requires_grad_(True)
    T = S if same else user_order_points(tree, "targets").clone().requires_grad_(True)
    pot = fmm_potential(actx, wr, actx.from_numpy(q), sources=S, targets=T)
    (torch.conj(actx.from_numpy(w))*pot).real.sum().backward()
    gS, gT = actx.to_numpy(S.grad), actx.to_numpy(T.grad)

    def L():
        p, = drive_fmm(actx, wr, [actx.from_numpy(q)])
        return float(np.real(np.vdot(w, actx.to_numpy(p))))
    h = 1e-6
    checks = [("sources", gS)] + ([] if same else [("targets", gT)])
    for which, grad in checks:
        pts = getattr(tree, which)        # device, tree order (moved in place)
        pos = _tree_index(tree, which)
        fd, got = [], []
        for _ in range(64 // len(checks)):
            a, u = int(rng.integers(0, 3)), int(rng.integers(0, pts.shape[1]))
            x0 = float(pts[a, pos[u]])
            pts[a, pos[u]] = x0 + h
            lp = L()
            pts[a, pos[u]] = x0 - h
            lm = L()
            pts[a, pos[u]] = x0
            fd.append((lp - lm)/(2*h))
            got.append(grad[a, u])
        fd, got = np.array(fd), np.array(got)
        # central differences err by ~(h/r)^2 for a point whose nearest
        # neighbour is r away (r >~ 1e-3 here)
        assert np.linalg.norm(got - fd)/np.linalg.norm(fd) < 1e-5, which


def test_fmm_source_gradient_matches_source_derivative_fmm(actx):
    """The source-position gradient is q_j (D_a^T g)_j with D_a the matrix of
    the AxisSourceDerivative(a) FMM (reference ``kernel.py:2288-2327``): for
    random w, sum_j w_j grad_s[a, j] / q_j equals <g, FMM_{d/dy_a}(w)>; the
    target-position gradient is g_t times the AxisTargetDerivative(a) FMM."""
    from sumpy_amd.autograd import fmm_potential
    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.kernel import AxisSourceDerivative, AxisTargetDerivative
    from sumpy_amd.tree import user_order_points
    rng, src, tgt = _separate(23)
    knl, tree, trav, wr = _fmm_setup(actx, "laplace", src, tgt, 8)
    ns, nt = tree.nsources, tree.ntargets
    q = 1.0 + rng.random(ns)
    g = rng.standard_normal(nt)
    S = user_order_points(tree, "sources").clone().requires_grad_(True)
    T = user_order_points(tree, "targets").clone().requires_grad_(True)
    pot = fmm_potential(actx, wr, actx.from_numpy(q), sources=S, targets=T)
    (actx.from_numpy(g)*pot).sum().backward()
    gS, gT = actx.to_numpy(S.grad), actx.to_numpy(T.grad)
    w = rng.standard_normal(ns)
    for a in range(3):
        _, _, _, wa = _fmm_setup(actx, "laplace", src, tgt, 8,
                                 source_kernels=[AxisSourceDerivative(a, knl)],
                                 tree_trav=(tree, trav))
        fa, = drive_fmm(actx, wa, [actx.from_numpy(w)])
        lhs = float(np.sum(w*gS[a]/q))
        rhs = float(g @ actx.to_numpy(fa))
        assert abs(lhs - rhs) <= 1e-12*(abs(rhs) + np.linalg.norm(g)*np.linalg.norm(
            actx.to_numpy(fa))), (a, lhs, rhs)
        _, _, _, ta = _fmm_setup(actx, "laplace", src, tgt, 8,
                                 target_kernels=[AxisTargetDerivative(a, knl)],
                                 tree_trav=(tree, trav))
        da, = drive_fmm(actx, ta, [actx.from_numpy(q)])
        assert np.linalg.norm(gT[a] - g*actx.to_numpy(da)) <= \
            1e-13*np.linalg.norm(g*actx.to_numpy(da))
