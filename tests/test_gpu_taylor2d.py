"""GPU parity of the 2D volume-Taylor expansions (Laplace and Helmholtz, full
and PDE-compressed; ``csrc/taylor2d.hip``) against the oracle
(``oracle/taylor2d.py``) operator by operator, and the 2D FMM against the
direct sum: the reference's ``test_sumpy_fmm`` cases with 2D kernels
(``sumpy/test/test_fmm.py:78-299``: 1000 normal sources, 300 targets offset
by 0.1, 30 per box, orders 1-3, k = 0.05, with and without translation
classes; error decreasing with the order) and its exclude-self case
(``test_fmm.py:538-598``: 500 sources, full Taylor order 10, rel. error
< 1e-7).
"""

from functools import partial

import numpy as np
import pytest
import torch

from oracle import taylor2d as t2

pytestmark = pytest.mark.gpu

K = 0.05
KC = 0.2*(0.707 + 0.707j)   # the reference's evanescent k (sumpy/test/test_kernels.py:168)


def _k(kind):
    return KC if kind == "helmc" else K


def _okind(kind):
    return "helm" if kind.startswith("helm") else "lap"


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b))/np.linalg.norm(np.asarray(b)))


def _classes(kind, compressed):
    from sumpy_amd.expansion.local import (
        LinearPDEConformingVolumeTaylorLocalExpansion,
        VolumeTaylorLocalExpansion,
    )
    from sumpy_amd.expansion.multipole import (
        LinearPDEConformingVolumeTaylorMultipoleExpansion,
        VolumeTaylorMultipoleExpansion,
    )
    from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel
    knl = LaplaceKernel(2) if kind == "lap" else HelmholtzKernel(
        2, allow_evanescent=kind == "helmc")
    if compressed:
        return (knl, LinearPDEConformingVolumeTaylorMultipoleExpansion,
                LinearPDEConformingVolumeTaylorLocalExpansion)
    return knl, VolumeTaylorMultipoleExpansion, VolumeTaylorLocalExpansion


def _chain_geometry():
    rng = np.random.default_rng(18)
    origin = np.array([2.0, 1.0])
    src = 0.7*(-0.5 + rng.random((2, 15))) + origin[:, None]
    q = np.ones(15)/15
    g = np.linspace(-0.15, 0.15, 6)
    off = np.array([5.5, 0.0])
    tx, ty = np.meshgrid(g + off[0] + origin[0], g + off[1] + origin[1], indexing="ij")
    tgt = np.array([tx.ravel(), ty.ravel()])
    c = np.array([[0, 0], [-0.2, 0.1], off + [0.3, -0.2], off]).T + origin[:, None]
    return src, q, tgt, c


@pytest.mark.parametrize("kind", ["lap", "helm", "helmc"])
@pytest.mark.parametrize("compressed", [True, False])
@pytest.mark.parametrize("order", [3, 7])
def test_taylor2d_operator_chain_matches_oracle(actx, kind, compressed, order):
    """Boxes 0..3 = the four centres; box 0 holds the sources, box 3 the
    targets; P2M(0) -> M2M(0->1) -> M2L(1->2, with and without class data)
    -> L2L(2->3) -> L2P(3), plus M2P(1) and P2L(3), through the operators'
    public calls."""
    from sumpy_amd import (
        E2EFromChildren,
        E2EFromCSR,
        E2EFromParent,
        E2PFromCSR,
        E2PFromSingleBox,
        M2LGenerateTranslationClassesDependentData,
        M2LUsingTranslationClassesDependentData,
        P2EFromCSR,
        P2EFromSingleBox,
    )
    knl, mcls, lcls = _classes(kind, compressed)
    ok, kv = _okind(kind), _k(kind)
    mpole, local = mcls(knl, order), lcls(knl, order)
    kw = {"k": _k(kind)} if kind != "lap" else {}
    cdt = torch.complex128 if kind != "lap" else torch.float64
    src, q, tgt, c = _chain_geometry()
    rs = (0.5, 0.25, 0.25, 0.125)
    dev = actx.from_numpy
    i32 = lambda a: dev(np.asarray(a, np.int32))  # noqa: E731
    nb = 4
    centers = dev(np.ascontiguousarray(c))
    srcd, tgtd = dev(src), dev(tgt)
    qd = dev(q.astype(np.complex128 if kind != "lap" else np.float64))
    bss, bsc = i32([0, 15, 15, 15]), i32([15, 0, 0, 0])
    bts, btc = i32([0, 0, 0, 0]), i32([0, 0, 0, tgt.shape[1]])
    ncoef = len(mpole)

    def zeros():
        return torch.zeros((nb, ncoef), dtype=cdt, device=actx.device)

    M = zeros()
    P2EFromSingleBox(mpole)(actx, sources=srcd, centers=centers, strengths=[qd],
                            tgt_expansions=M, tgt_base_ibox=0, rscale=rs[0],
                            source_boxes=i32([0]), box_source_starts=bss,
                            box_source_counts_nonchild=bsc, **kw)
    child = np.zeros((4, nb), np.int32)
    child[0, 1] = 2
    Msrc = zeros()
    Msrc[2] = M[0]
    cen2 = c.copy()
    cen2[:, 2] = c[:, 0]
    M2 = zeros()
    E2EFromChildren(mpole, mpole)(actx, src_expansions=Msrc, src_base_ibox=0,
                                  tgt_expansions=M2, tgt_base_ibox=0, target_boxes=i32([1]),
                                  box_child_ids=dev(child),
                                  centers=dev(np.ascontiguousarray(cen2)),
                                  src_rscale=rs[0], tgt_rscale=rs[1], **kw)
    L = zeros()
    E2EFromCSR(mpole, local)(actx, src_expansions=M2, src_base_ibox=0, tgt_expansions=L,
                             tgt_base_ibox=0, target_boxes=i32([2]), src_box_starts=i32([0, 1]),
                             src_box_lists=i32([1]), centers=centers, src_rscale=rs[1],
                             tgt_rscale=rs[2], **kw)
    # the same M2L through the translation-class operators (one class)
    vec = dev(np.ascontiguousarray((c[:, 2] - c[:, 1])[:, None]))
    nd = local.m2l_translation.translation_classes_dependent_ndata(local, mpole)
    data = torch.zeros((1, nd), dtype=torch.float64, device=actx.device)
    M2LGenerateTranslationClassesDependentData(mpole, local)(
        actx, src_rscale=rs[1], translation_classes_level_start=0, ntranslation_classes=1,
        m2l_translation_classes_dependent_data=data, m2l_translation_vectors=vec,
        ntranslation_vectors=1, **kw)
    Lc = zeros()
    M2LUsingTranslationClassesDependentData(mpole, local)(
        actx, src_expansions=M2, src_base_ibox=0, tgt_expansions=Lc, tgt_base_ibox=0,
        target_boxes=i32([2]), src_box_starts=i32([0, 1]), src_box_lists=i32([1]),
        centers=centers, src_rscale=rs[1], tgt_rscale=rs[2],
        m2l_translation_classes_dependent_data=data, translation_classes_level_start=0,
        m2l_translation_classes_lists=i32([0]), **kw)
    L2 = zeros()
    E2EFromParent(local, local)(actx, src_expansions=L, src_base_ibox=0, tgt_expansions=L2,
                                tgt_base_ibox=0, target_boxes=i32([3]),
                                box_parent_ids=i32([0, 0, 0, 2]), centers=centers,
                                src_rscale=rs[2], tgt_rscale=rs[3], **kw)
    pot_l2p, = E2PFromSingleBox(local, [knl])(
        actx, src_expansions=L2, src_base_ibox=0, target_boxes=i32([3]), centers=centers,
        targets=tgtd, box_target_starts=bts, box_target_counts_nonchild=btc, rscale=rs[3],
        **kw)
    pot_m2p, = E2PFromCSR(mpole, [knl])(
        actx, src_expansions=M2, src_base_ibox=0, target_boxes=i32([3]),
        source_box_starts=i32([0, 1]), source_box_lists=i32([1]), centers=centers,
        targets=tgtd, box_target_starts=bts, box_target_counts_nonchild=btc, rscale=rs[1],
        **kw)
    L4 = zeros()
    P2EFromCSR(local)(actx, sources=srcd, centers=centers, strengths=[qd], tgt_expansions=L4,
                      tgt_base_ibox=0, rscale=rs[3], target_boxes=i32([3]),
                      source_box_starts=i32([0, 1]), source_box_lists=i32([0]),
                      box_source_starts=bss, box_source_counts_nonchild=bsc, **kw)

    m1 = t2.p2m(order, compressed, ok, kv, c[:, 0], src, q, rs[0])
    m2 = t2.m2m(order, order, compressed, ok, kv, m1, c[:, 0], c[:, 1], rs[0], rs[1])
    l1 = t2.m2l(order, order, compressed, ok, kv, m2, c[:, 2] - c[:, 1], rs[1])
    l2 = t2.l2l(order, order, compressed, ok, kv, l1, c[:, 2], c[:, 3], rs[2], rs[3])
    l4 = t2.p2l(order, compressed, ok, kv, c[:, 3], src, q, rs[3])
    got = {n: actx.to_numpy(v) for n, v in (("p2m", M[0]), ("m2m", M2[1]), ("m2l", L[2]),
                                             ("m2l_classes", Lc[2]), ("l2l", L2[3]),
                                             ("p2l", L4[3]))}
    want = {"p2m": m1, "m2m": m2, "m2l": l1, "m2l_classes": l1, "l2l": l2, "p2l": l4}
    for n in want:
        e = rel(got[n], want[n])
        print(f"{kind} compressed={compressed} p={order} {n}: {e:.2e}")
        assert e < 1e-12, n
    ref = t2.l2p(order, compressed, ok, kv, c[:, 3], l2, tgt, rs[3])
    assert rel(actx.to_numpy(pot_l2p), ref) < 1e-12
    ref = t2.m2p(order, compressed, ok, kv, c[:, 1], m2, tgt, rs[1])
    assert rel(actx.to_numpy(pot_m2p), ref) < 1e-12
    # (the chain itself converges to the direct sum)
    assert rel(actx.to_numpy(pot_l2p), t2.p2p(ok, kv, tgt, src, q)) < 0.2**order


def _fmm(actx, tree, trav, q, kind, compressed, order, use_classes, exclude_self=False):
    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    knl, mcls, lcls = _classes(kind, compressed)
    ti = SumpyTreeIndependentDataForWrangler(actx, partial(mcls, knl), partial(lcls, knl),
                                             [knl], exclude_self=exclude_self)
    self_kw = {}
    if exclude_self:
        self_kw["target_to_source"] = actx.from_numpy(np.arange(tree.ntargets, dtype=np.int32))
    dtype = np.complex128 if kind != "lap" else np.float64
    wr = SumpyExpansionWrangler(ti, trav, dtype, lambda *a: order,
                                kernel_extra_kwargs={"k": _k(kind)} if kind != "lap" else None,
                                self_extra_kwargs=self_kw,
                                _disable_translation_classes=not use_classes)
    print(f"fmm {kind} compressed={compressed} p={order} classes={use_classes}", flush=True)
    pot, = drive_fmm(actx, wr, [actx.from_numpy(q.astype(dtype))])
    return actx.to_numpy(pot)


@pytest.mark.parametrize("kind", ["lap", "helm", "helmc"])
@pytest.mark.parametrize("compressed", [True, False])
@pytest.mark.parametrize("use_classes", [True, False])
def test_taylor2d_fmm_converges(actx, kind, compressed, use_classes):
    from sumpy_amd.tree import build_tree_and_traversal
    src = np.random.default_rng(15).standard_normal((2, 1000))
    tgt = np.array([[0.1], [0.0]]) + np.random.default_rng(18).standard_normal((2, 300))
    q = np.random.default_rng(44).random(1000)
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    direct = t2.p2p(_okind(kind), _k(kind), tgt, src, q)
    errs = []
    for order in (1, 2, 3):
        pot = _fmm(actx, tree, trav, q, kind, compressed, order, use_classes)
        errs.append(float(np.max(np.abs(pot - direct))/np.max(np.abs(direct))))
    print(f"{kind} compressed={compressed} classes={use_classes}: rel-inf errors {errs}")
    assert errs[2] < errs[1] < errs[0] < 0.1
    # order of convergence: about one digit per order at 30 per box
    assert errs[2] < 0.05*errs[0]


@pytest.mark.parametrize("compressed", [True, False])
def test_taylor2d_fmm_exclude_self(actx, compressed):
    from sumpy_amd.tree import build_tree_and_traversal
    src = np.random.default_rng(15).standard_normal((2, 500))
    q = np.random.default_rng(44).random(500)
    tree, trav = build_tree_and_traversal(actx, src, None, 30)
    pot = _fmm(actx, tree, trav, q, "lap", compressed, 10, True, exclude_self=True)
    r = np.hypot(src[0][:, None] - src[0][None, :], src[1][:, None] - src[1][None, :])
    np.fill_diagonal(r, 1.0)
    direct = -1/(2*np.pi)*(np.log(r) @ q)
    e = rel(pot, direct)
    print(f"exclude self, order 10, compressed={compressed}: rel. l2 error {e:.2e}")
    assert e < 1e-7


def _der_kernels(kind, compressed):
    from sumpy_amd.kernel import AxisTargetDerivative, DirectionalSourceDerivative
    knl, mcls, lcls = _classes(kind, compressed)
    skn = [knl, DirectionalSourceDerivative(knl, "dir_vec")]
    tkn = [knl, AxisTargetDerivative(0, knl), AxisTargetDerivative(1, knl)]
    return knl, mcls, lcls, skn, tkn


@pytest.mark.parametrize("kind", ["lap", "helm", "helmc"])
@pytest.mark.parametrize("compressed", [True, False])
def test_taylor2d_dipole_p2e_and_gradient_e2p_match_oracle(actx, kind, compressed):
    """P2M and P2L with charge + DirectionalSourceDerivative sources and M2P /
    L2P with potential + d/dx + d/dy targets (the kernels of the reference's
    test_p2e2p with a source derivative), through the operators, against
    the oracle."""
    from sumpy_amd import E2PFromCSR, E2PFromSingleBox, P2EFromCSR, P2EFromSingleBox
    knl, mcls, lcls, skn, tkn = _der_kernels(kind, compressed)
    ok, kv = _okind(kind), _k(kind)
    order = 6
    mpole, local = mcls(knl, order), lcls(knl, order)
    kw = {"k": kv} if kind != "lap" else {}
    cx = kind != "lap"
    src, _, tgt, c = _chain_geometry()
    rng = np.random.default_rng(3)
    q0, q1 = rng.standard_normal(15), rng.standard_normal(15)
    n = rng.standard_normal((2, 15))
    dip = (q1*n[0], q1*n[1])
    d = actx.from_numpy
    i32 = lambda a: d(np.asarray(a, np.int32))  # noqa: E731
    wdt = np.complex128 if cx else np.float64
    cdt = torch.complex128 if cx else torch.float64
    centers = d(np.ascontiguousarray(c))
    bss, bsc = i32([0, 15, 15, 15]), i32([15, 0, 0, 0])
    bts, btc = i32([0, 0, 0, 0]), i32([0, 0, 0, tgt.shape[1]])
    nst = len(t2.stored_ids(order, compressed, ok))
    print(f"taylor2d dipole/gradient operators {kind} compressed={compressed}", flush=True)
    M = torch.zeros((4, nst), dtype=cdt, device=actx.device)
    P2EFromSingleBox(mpole, kernels=skn, strength_usage=[0, 1])(
        actx, sources=d(src), centers=centers, strengths=[d(q0.astype(wdt)), d(q1.astype(wdt))],
        dir_vec=d(n), tgt_expansions=M, tgt_base_ibox=0, rscale=0.5, source_boxes=i32([0]),
        box_source_starts=bss, box_source_counts_nonchild=bsc, **kw)
    m_ref = t2.p2m(order, compressed, ok, kv, c[:, 0], src, q0, 0.5, dipoles=dip)
    assert rel(actx.to_numpy(M[0]), m_ref) < 1e-12
    L = torch.zeros((4, nst), dtype=cdt, device=actx.device)
    P2EFromCSR(local, kernels=skn, strength_usage=[0, 1])(
        actx, sources=d(src), centers=centers, strengths=[d(q0.astype(wdt)), d(q1.astype(wdt))],
        dir_vec=d(n), tgt_expansions=L, tgt_base_ibox=0, rscale=0.25, target_boxes=i32([3]),
        source_box_starts=i32([0, 1]), source_box_lists=i32([0]), box_source_starts=bss,
        box_source_counts_nonchild=bsc, **kw)
    l_ref = t2.p2l(order, compressed, ok, kv, c[:, 3], src, q0, 0.25, dipoles=dip)
    assert rel(actx.to_numpy(L[3]), l_ref) < 1e-12
    got = E2PFromSingleBox(local, tkn)(
        actx, src_expansions=L, src_base_ibox=0, target_boxes=i32([3]), centers=centers,
        targets=d(tgt), box_target_starts=bts, box_target_counts_nonchild=btc, rscale=0.25,
        **kw)
    ref = t2.l2p(order, compressed, ok, kv, c[:, 3], l_ref, tgt, 0.25, grad=True)
    for g, r in zip(got, ref):
        assert rel(actx.to_numpy(g), r) < 1e-12
    Mb = torch.zeros_like(M)
    Mb[1] = M[0]
    cen = c.copy()
    cen[:, 1] = c[:, 0]
    got = E2PFromCSR(mpole, tkn)(
        actx, src_expansions=Mb, src_base_ibox=0, target_boxes=i32([3]),
        source_box_starts=i32([0, 1]), source_box_lists=i32([1]),
        centers=d(np.ascontiguousarray(cen)), targets=d(tgt), box_target_starts=bts,
        box_target_counts_nonchild=btc, rscale=0.5, **kw)
    ref = t2.m2p(order, compressed, ok, kv, c[:, 0], m_ref, tgt, 0.5, grad=True)
    for g, r in zip(got, ref):
        assert rel(actx.to_numpy(g), r) < 1e-12
    direct = t2.p2p(ok, kv, tgt, src, q0, dipoles=dip, grad=True)
    for g, r in zip(got, direct):
        assert rel(actx.to_numpy(g), r) < 1e-4


@pytest.mark.parametrize("kind", ["lap", "helm"])
def test_taylor2d_fmm_dipoles_and_gradients(actx, kind):
    """The 2D volume FMM (compressed) with charge + dipole sources and
    potential + gradient targets against the direct sum, converging with the
    order."""
    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.tree import build_tree_and_traversal
    knl, mcls, lcls, skn, tkn = _der_kernels(kind, True)
    src = np.random.default_rng(15).standard_normal((2, 1000))
    tgt = np.array([[0.1], [0.0]]) + np.random.default_rng(18).standard_normal((2, 300))
    q = np.random.default_rng(44).random(1000)
    rng = np.random.default_rng(21)
    q1 = rng.standard_normal(1000)
    n = rng.standard_normal((2, 1000))
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    ids = tree.host["user_source_ids"]
    ref = t2.p2p(kind, K, tgt, src, q, dipoles=(q1*n[0], q1*n[1]), grad=True)
    dtype = np.complex128 if kind != "lap" else np.float64
    errs = []
    for order in (3, 6):
        ti = SumpyTreeIndependentDataForWrangler(actx, partial(mcls, knl), partial(lcls, knl),
                                                 tkn, source_kernels=skn, strength_usage=[0, 1])
        wr = SumpyExpansionWrangler(ti, trav, dtype, lambda *a: order,
                                    kernel_extra_kwargs={"k": K} if kind != "lap" else None,
                                    source_extra_kwargs={"dir_vec": actx.from_numpy(
                                        np.ascontiguousarray(n[:, ids]))})
        print(f"taylor2d fmm derivatives {kind} p={order}", flush=True)
        pots = drive_fmm(actx, wr, [actx.from_numpy(q.astype(dtype)),
                                    actx.from_numpy(q1.astype(dtype))])
        errs.append(max(rel(actx.to_numpy(p), r) for p, r in zip(pots, ref)))
    print(f"{kind}: dipole/gradient FMM errors {errs}")
    assert errs[1] < 0.1*errs[0] and errs[0] < 1e-2


@pytest.mark.parametrize("local", [True, False])
def test_p2e_multiple_source_kernels_is_linear(actx, local):
    """The reference's ``test_p2e_multiple`` (``sumpy/test/test_kernels.py:152-260``):
    2D Laplace, PDE-compressed order 4, 100 sources, P2EFromSingleBox with
    [DirectionalSourceDerivative, charge] and strength_usage [0, 1] in one
    call (tgt_expansions made from nboxes) equals the sum of the one-kernel
    calls to 1e-12."""
    from sumpy_amd import P2EFromSingleBox
    from sumpy_amd.kernel import DirectionalSourceDerivative
    knl, mcls, lcls = _classes("lap", True)
    expn = (lcls if local else mcls)(knl, 4)
    skn = [DirectionalSourceDerivative(knl, "dir_vec"), knl]
    n = 100
    rng = np.random.default_rng(14)
    center = np.array([2.0, 1.0])
    d = actx.from_numpy
    src = d(0.7*(-0.5 + rng.random((2, n))) + center[:, None])
    strengths = [d(np.full(n, 1/n)), d(np.full(n, 2/n))]
    alpha = np.linspace(0, 2*np.pi, n)
    dir_vec = d(np.vstack([np.cos(alpha), np.sin(alpha)]))
    c = (np.array([5.5, 0.0]) + center) if local else center
    kw = dict(source_boxes=d(np.array([0], np.int32)),
              box_source_starts=d(np.array([0], np.int32)),
              box_source_counts_nonchild=d(np.array([n], np.int32)),
              centers=d(c.reshape(2, 1)), sources=src, strengths=strengths, nboxes=1,
              tgt_base_ibox=0, rscale=0.5)
    actual = actx.to_numpy(P2EFromSingleBox(expn, kernels=skn, strength_usage=[0, 1])(
        actx, dir_vec=dir_vec, **kw))
    expected = sum(
        actx.to_numpy(P2EFromSingleBox(expn, kernels=[sk], strength_usage=[i])(
            actx, **kw, **({"dir_vec": dir_vec} if i == 0 else {})))
        for i, sk in enumerate(skn))
    assert np.linalg.norm(actual - expected)/np.linalg.norm(expected) < 1e-12


def _fmm_2d_laplace(actx, tree, trav, q, order, target_kernels, source_kernels):
    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.expansion.local import VolumeTaylorLocalExpansion
    from sumpy_amd.expansion.multipole import VolumeTaylorMultipoleExpansion
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.kernel import LaplaceKernel
    knl = LaplaceKernel(2)
    ti = SumpyTreeIndependentDataForWrangler(
        actx, partial(VolumeTaylorMultipoleExpansion, knl),
        partial(VolumeTaylorLocalExpansion, knl), target_kernels=target_kernels,
        source_kernels=source_kernels, exclude_self=True)
    wr = SumpyExpansionWrangler(
        ti, trav, np.float64, lambda *a: order,
        self_extra_kwargs={"target_to_source": actx.from_numpy(
            np.arange(tree.ntargets, dtype=np.int32))})
    return [actx.to_numpy(p) for p in drive_fmm(actx, wr, [actx.from_numpy(q)])]


def test_fmm_axis_source_derivative_2d(actx):
    """The reference's ``test_sumpy_axis_source_derivative``
    (``sumpy/test/test_fmm.py:605-665``): 2D Laplace, full Taylor order 10,
    500 normal sources = targets (self excluded), 30 per box; the FMM of
    AxisTargetDerivative(0) with charges equals minus the FMM of
    AxisSourceDerivative(0) (relative l2 < 1e-5)."""
    from sumpy_amd.kernel import AxisSourceDerivative, AxisTargetDerivative, LaplaceKernel
    from sumpy_amd.tree import build_tree_and_traversal
    knl = LaplaceKernel(2)
    src = np.random.default_rng(15).standard_normal((2, 500))
    q = np.random.default_rng(12).random(500)
    tree, trav = build_tree_and_traversal(actx, src, None, 30)
    p0, = _fmm_2d_laplace(actx, tree, trav, q, 10, [AxisTargetDerivative(0, knl)], [knl])
    p1, = _fmm_2d_laplace(actx, tree, trav, q, 10, [knl], [AxisSourceDerivative(0, knl)])
    e = np.linalg.norm(p0 + p1)/np.linalg.norm(p0)
    print(f"axis source derivative 2D: {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("deriv_axes", [(), (0,), (1,)])
def test_fmm_target_point_multiplier_2d(actx, deriv_axes):
    """The reference's ``test_sumpy_target_point_multiplier``
    (``sumpy/test/test_fmm.py:672-741``): 2D Laplace, full Taylor order 5;
    [x0 G, G, G] with the derivatives of deriv_axes applied to the first two:
    x0 pot (+ pot for d/dx0) equals the first (relative l2 < 1e-5)."""
    from sumpy_amd.kernel import AxisTargetDerivative, LaplaceKernel, TargetPointMultiplier
    from sumpy_amd.tree import build_tree_and_traversal
    knl = LaplaceKernel(2)
    src = np.random.default_rng(15).standard_normal((2, 500))
    q = np.random.default_rng(12).random(500)
    tree, trav = build_tree_and_traversal(actx, src, None, 30)
    tk = [TargetPointMultiplier(0, knl), knl, knl]
    for axis in deriv_axes:
        tk[0] = AxisTargetDerivative(axis, tk[0])
        tk[1] = AxisTargetDerivative(axis, tk[1])
    p0, p1, p2 = _fmm_2d_laplace(actx, tree, trav, q, 5, tk, [knl])
    ref = p1*src[0] + (p2 if deriv_axes == (0,) else 0)
    e = np.linalg.norm(p0 - ref)/np.linalg.norm(ref)
    print(f"target point multiplier 2D {deriv_axes}: {e:.2e}")
    assert e < 1e-5
