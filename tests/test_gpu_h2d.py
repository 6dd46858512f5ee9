"""2D Helmholtz and Yukawa (H2D / Y2D Hankel/Bessel expansions) on the GPU
against the oracle.  The Yukawa cases run with the Bessel argument i lam in
the oracle (scipy at complex arguments) and through I_n, K_n on the device.

* The dense P2P equals ``oracle.h2d.p2p`` (scipy's H0) to 1e-13 relative,
  also at k |d| up to ~60 (the device's Miller/Neumann Bessel routine
  against scipy).
* The expansion chains of the reference's ``test_translations``
  (``sumpy/test/test_kernels.py:493-640``) through the wrangler's operators
  equal the oracle's chains to 1e-12.
* The full FMM on the reference's ``test_sumpy_fmm`` H2D setup
  (``sumpy/test/test_fmm.py:78-299``: 1000 normal sources, 300 targets
  offset by 0.1, weights ``default_rng(44).random``, 30 per box, k = 0.05,
  orders 4 and 5, with and without translation classes) equals the
  oracle-driven FMM (``tests/oracle_fmm.py:oracle_fmm_h2d``) to 1e-12 and
  converges to the direct sum.
"""

import numpy as np
import pytest
import torch

from oracle import h2d as oh
from oracle_fmm import oracle_fmm_h2d

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b)/np.linalg.norm(b))


def _knl():
    from sumpy_amd.kernel import HelmholtzKernel
    return HelmholtzKernel(2)


def _kernel(kind):
    """(kernel, multipole class, local class, keyword name, oracle argument
    scale from the parameter)."""
    from sumpy_amd.expansion.local import H2DLocalExpansion, Y2DLocalExpansion
    from sumpy_amd.expansion.multipole import H2DMultipoleExpansion, Y2DMultipoleExpansion
    from sumpy_amd.kernel import HelmholtzKernel, YukawaKernel
    if kind in ("helmholtz", "helmholtz_c"):
        return (HelmholtzKernel(2, allow_evanescent=kind == "helmholtz_c"),
                H2DMultipoleExpansion, H2DLocalExpansion, "k", lambda v: v)
    return YukawaKernel(2), Y2DMultipoleExpansion, Y2DLocalExpansion, "lam", lambda v: 1j*v


@pytest.mark.parametrize("lam", [0.05, 2.0, 20.0])
def test_y2d_p2p_dense_matches_oracle(actx, lam):
    from sumpy_amd import P2P
    knl = _kernel("yukawa")[0]
    rng = np.random.default_rng(8)
    src, tgt = 4*rng.random((2, 700)), 4*rng.random((2, 300))
    q = rng.standard_normal(700) + 1j*rng.standard_normal(700)
    pot, = P2P([knl], exclude_self=False)(actx, actx.from_numpy(tgt), actx.from_numpy(src),
                                           [actx.from_numpy(q)], lam=lam)
    err = rel(actx.to_numpy(pot), oh.p2p(1j*lam, tgt, src, q))
    print(f"lam={lam}: Yukawa P2P vs oracle {err:.2e}")
    assert err < 1e-13


@pytest.mark.parametrize("k", [0.05, 1.0, 15.0])
def test_h2d_p2p_dense_matches_oracle(actx, k):
    from sumpy_amd import P2P
    rng = np.random.default_rng(5)
    src, tgt = 4*rng.random((2, 700)), 4*rng.random((2, 300))
    q = rng.standard_normal(700) + 1j*rng.standard_normal(700)
    pot, = P2P([_knl()], exclude_self=False)(actx, actx.from_numpy(tgt), actx.from_numpy(src),
                                              [actx.from_numpy(q)], k=k)
    err = rel(actx.to_numpy(pot), oh.p2p(k, tgt, src, q))
    print(f"k={k}: P2P vs oracle {err:.2e}")
    assert err < 1e-13


@pytest.mark.parametrize("k", [0.2*(0.707 + 0.707j), 3.0 + 1.0j, 2.0j])
def test_h2d_p2p_dense_complex_k_matches_oracle(actx, k):
    """Damped waves: H0 of complex argument (J + i Y by the complex
    recurrences); the first k is the reference's evanescent value
    (``sumpy/test/test_kernels.py:168``)."""
    from sumpy_amd import P2P
    rng = np.random.default_rng(5)
    src, tgt = 4*rng.random((2, 700)), 4*rng.random((2, 300))
    q = rng.standard_normal(700) + 1j*rng.standard_normal(700)
    pot, = P2P([_kernel("helmholtz_c")[0]], exclude_self=False)(
        actx, actx.from_numpy(tgt), actx.from_numpy(src), [actx.from_numpy(q)], k=k)
    err = rel(actx.to_numpy(pot), oh.p2p(k, tgt, src, q))
    print(f"k={k}: P2P vs oracle {err:.2e}")
    # H0 = J0 + i Y0 cancels: |J0|, |Y0| ~ e^{Im k r} against |H0| ~ e^{-Im k r},
    # so the relative accuracy of a pair is about 1e-16 e^{2 Im k r}
    # (k = 2i, r up to 5.7 here)
    assert err < (1e-13 if k.imag < 1.5 else 2e-12)
    with pytest.raises(ValueError, match="allow_evanescent"):
        P2P([_knl()], exclude_self=False)(
            actx, actx.from_numpy(tgt), actx.from_numpy(src), [actx.from_numpy(q)], k=k)


def test_h2d_p2p_exclude_self(actx):
    from sumpy_amd import P2P
    rng = np.random.default_rng(6)
    pts = rng.random((2, 400))
    q = rng.random(400)
    t2s = np.arange(400, dtype=np.int32)
    pot, = P2P([_knl()], exclude_self=True)(
        actx, actx.from_numpy(pts), actx.from_numpy(pts), [actx.from_numpy(q)], k=2.0,
        target_to_source=actx.from_numpy(t2s))
    assert rel(actx.to_numpy(pot), oh.p2p(2.0, pts, pts, q, exclude_self_t2s=t2s)) < 1e-13


def _chain_geometry():
    rng = np.random.default_rng(18)
    origin = np.array([2.0, 1.0])
    src = 0.7*(-0.5 + rng.random((2, 15))) + origin[:, None]
    g = np.linspace(-0.15, 0.15, 20)
    off = np.array([5.5, 0.0])
    tx, ty = np.meshgrid(g + off[0] + origin[0], g + off[1] + origin[1], indexing="ij")
    tgt = np.array([tx.ravel(), ty.ravel()])
    centers = np.array([[0, 0], [-0.2, 0.1], off + [0.3, -0.2], off]).T + origin[:, None]
    return src, np.ones(15)/15, tgt, centers


@pytest.mark.parametrize(("kind", "order", "k"), [
    ("helmholtz", 3, 0.05), ("helmholtz", 8, 0.05), ("helmholtz", 20, 1.3),
    ("helmholtz_c", 8, 0.2*(0.707 + 0.707j)), ("helmholtz_c", 20, 1.3 + 0.4j),
    ("yukawa", 8, 2.0), ("yukawa", 20, 0.5)])
def test_h2d_operator_chain_matches_oracle(actx, kind, order, k):
    """Boxes 0..3 = the four centres; box 0 holds the sources, box 3 the
    targets; P2M(0) -> M2M(0->1) -> M2L(1->2) -> L2L(2->3) -> L2P(3), plus
    M2P(1) and P2L(3), each through the operators' public calls."""
    from sumpy_amd import (
        E2EFromChildren,
        E2EFromCSR,
        E2EFromParent,
        E2PFromCSR,
        E2PFromSingleBox,
        P2EFromCSR,
        P2EFromSingleBox,
    )
    knl, mcls, lcls, kname, arg = _kernel(kind)
    mpole, local = mcls(knl, order), lcls(knl, order)
    kw = {kname: k}
    ka = arg(k)
    src, q, tgt, c = _chain_geometry()
    rs = (0.5, 0.25, 0.5, 0.25)
    dev = actx.from_numpy
    i32 = lambda a: dev(np.asarray(a, np.int32))  # noqa: E731
    nb = 4
    centers = dev(np.ascontiguousarray(c))
    srcd, tgtd, qd = dev(src), dev(tgt), dev(q.astype(np.complex128))
    bss, bsc = i32([0, 15, 15, 15]), i32([15, 0, 0, 0])
    bts, btc = i32([0, 0, 0, 0]), i32([0, 0, 0, 400])
    ncoef = 2*order + 1

    def zeros():
        return torch.zeros((nb, ncoef), dtype=torch.complex128, device=actx.device)

    M = zeros()
    P2EFromSingleBox(mpole)(actx, sources=srcd, centers=centers, strengths=[qd],
                            tgt_expansions=M, tgt_base_ibox=0, rscale=rs[0], source_boxes=i32([0]),
                            box_source_starts=bss, box_source_counts_nonchild=bsc, **kw)
    # M2M with box 1 as the parent of box 0: a child id 0 means "no child",
    # so the M2M input is a copy with box 0's multipole and centre in row 2
    child = np.zeros((4, nb), np.int32)
    child[0, 1] = 2
    Msrc = zeros()
    Msrc[2] = M[0]
    cen2 = c.copy()
    cen2[:, 2] = c[:, 0]
    M2 = zeros()
    E2EFromChildren(mpole, mpole)(actx, src_expansions=Msrc, src_base_ibox=0,
                                  tgt_expansions=M2, tgt_base_ibox=0, target_boxes=i32([1]),
                                  box_child_ids=dev(child), centers=dev(np.ascontiguousarray(cen2)),
                                  src_rscale=rs[0], tgt_rscale=rs[1], **kw)
    L = zeros()
    E2EFromCSR(mpole, local)(actx, src_expansions=M2, src_base_ibox=0, tgt_expansions=L,
                             tgt_base_ibox=0, target_boxes=i32([2]), src_box_starts=i32([0, 1]),
                             src_box_lists=i32([1]), centers=centers, src_rscale=rs[1],
                             tgt_rscale=rs[2], **kw)
    L2 = zeros()
    E2EFromParent(local, local)(actx, src_expansions=L, src_base_ibox=0, tgt_expansions=L2,
                                tgt_base_ibox=0, target_boxes=i32([3]),
                                box_parent_ids=i32([0, 0, 0, 2]), centers=centers,
                                src_rscale=rs[2], tgt_rscale=rs[3], **kw)
    pot_l2p, = E2PFromSingleBox(local, [knl])(
        actx, src_expansions=L2, src_base_ibox=0, target_boxes=i32([3]), centers=centers,
        targets=tgtd, box_target_starts=bts, box_target_counts_nonchild=btc, rscale=rs[3], **kw)
    pot_m2p, = E2PFromCSR(mpole, [knl])(
        actx, src_expansions=M2, src_base_ibox=0, target_boxes=i32([3]),
        source_box_starts=i32([0, 1]), source_box_lists=i32([1]), centers=centers,
        targets=tgtd, box_target_starts=bts, box_target_counts_nonchild=btc, rscale=rs[1], **kw)
    L4 = zeros()
    P2EFromCSR(local)(actx, sources=srcd, centers=centers, strengths=[qd], tgt_expansions=L4,
                      tgt_base_ibox=0, rscale=rs[3], target_boxes=i32([3]),
                      source_box_starts=i32([0, 1]), source_box_lists=i32([0]),
                      box_source_starts=bss, box_source_counts_nonchild=bsc, **kw)

    m1 = oh.p2m(order, ka, c[:, 0], src, q, rs[0])
    m2 = oh.m2m(order, order, ka, m1, c[:, 0], c[:, 1], rs[0], rs[1])
    l1 = oh.m2l(order, order, ka, m2, c[:, 2] - c[:, 1], rs[1], rs[2])
    l2 = oh.l2l(order, order, ka, l1, c[:, 2], c[:, 3], rs[2], rs[3])
    l4 = oh.p2l(order, ka, c[:, 3], src, q, rs[3])
    got = {n: actx.to_numpy(v) for n, v in (("p2m", M[0]), ("m2m", M2[1]), ("m2l", L[2]),
                                             ("l2l", L2[3]), ("p2l", L4[3]))}
    want = {"p2m": m1, "m2m": m2, "m2l": l1, "l2l": l2, "p2l": l4}
    for n in want:
        e = rel(got[n], want[n])
        print(f"{kind} p={order} k={k} {n}: {e:.2e}")
        assert e < 1e-12, n
    assert rel(actx.to_numpy(pot_l2p), oh.l2p(order, ka, c[:, 3], l2, tgt, rs[3])) < 1e-12
    assert rel(actx.to_numpy(pot_m2p), oh.m2p(order, ka, c[:, 1], m2, tgt, rs[1])) < 1e-12


def _fmm_case():
    src = np.random.default_rng(15).standard_normal((2, 1000))
    tgt = np.array([[0.1], [0.0]]) + np.random.default_rng(18).standard_normal((2, 300))
    q = np.random.default_rng(44).random(1000)
    return src, tgt, q


def _run_fmm(actx, tree, trav, q, order, k, use_classes, kind="helmholtz"):
    from functools import partial

    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    knl, mcls, lcls, kname, _ = _kernel(kind)
    ti = SumpyTreeIndependentDataForWrangler(actx, partial(mcls, knl), partial(lcls, knl),
                                             [knl])
    wr = SumpyExpansionWrangler(ti, trav, np.complex128, lambda *a: order,
                                kernel_extra_kwargs={kname: k},
                                _disable_translation_classes=not use_classes)
    pot, = drive_fmm(actx, wr, [actx.from_numpy(q.astype(np.complex128))])
    return actx.to_numpy(pot)


@pytest.mark.parametrize("use_classes", [True, False])
def test_h2d_fmm_matches_oracle_fmm_and_converges(actx, use_classes):
    from sumpy_amd.tree import build_tree_and_traversal
    src, tgt, q = _fmm_case()
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    direct = oh.p2p(0.05, tgt, src, q)
    errs = []
    for order in (4, 5):
        pot = _run_fmm(actx, tree, trav, q, order, 0.05, use_classes)
        ref = oracle_fmm_h2d(tree, trav, q, order, 0.05)
        e_or = rel(pot, ref)
        errs.append(float(np.max(np.abs(pot - direct))/np.max(np.abs(direct))))
        print(f"classes={use_classes} p={order}: vs oracle FMM {e_or:.2e}, "
              f"rel-inf vs direct {errs[-1]:.2e}")
        assert e_or < 1e-12
    assert errs[1] < errs[0] < 1e-3


def test_h2d_fmm_sources_are_targets_exclude_self_high_order(actx):
    """Sources = targets, self excluded, order 16, k = 6 (k h up to ~6 at
    the top levels), against the oracle FMM and the direct sum."""
    from functools import partial

    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.expansion.local import H2DLocalExpansion
    from sumpy_amd.expansion.multipole import H2DMultipoleExpansion
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.tree import build_tree_and_traversal
    rng = np.random.default_rng(2)
    pts = rng.random((2, 3000))
    q = rng.standard_normal(3000) + 1j*rng.standard_normal(3000)
    tree, trav = build_tree_and_traversal(actx, pts, None, 40)
    knl = _knl()
    ti = SumpyTreeIndependentDataForWrangler(actx, partial(H2DMultipoleExpansion, knl),
                                             partial(H2DLocalExpansion, knl), [knl],
                                             exclude_self=True)
    t2s = np.arange(3000, dtype=np.int32)
    wr = SumpyExpansionWrangler(ti, trav, np.complex128, lambda *a: 16,
                                kernel_extra_kwargs={"k": 6.0},
                                self_extra_kwargs={"target_to_source": actx.from_numpy(t2s)})
    pot, = drive_fmm(actx, wr, [actx.from_numpy(q)])
    pot = actx.to_numpy(pot)
    ref = oracle_fmm_h2d(tree, trav, q, 16, 6.0, exclude_self=True)
    direct = oh.p2p(6.0, pts, pts, q, exclude_self_t2s=t2s)
    print(f"p=16 k=6: vs oracle FMM {rel(pot, ref):.2e}, vs direct {rel(pot, direct):.2e}")
    assert rel(pot, ref) < 1e-11
    assert rel(pot, direct) < 1e-8


@pytest.mark.parametrize("use_classes", [True, False])
def test_y2d_fmm_matches_oracle_fmm_and_converges(actx, use_classes):
    """The reference's ``test_sumpy_fmm`` Y2D case (``sumpy/test/test_fmm.py:106-107,233-241``:
    lam = 2, orders 10 and 12) on the same particles as the H2D case."""
    from sumpy_amd.tree import build_tree_and_traversal
    src, tgt, q = _fmm_case()
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    direct = oh.p2p(2j, tgt, src, q)
    errs = []
    for order in (10, 12):
        pot = _run_fmm(actx, tree, trav, q, order, 2.0, use_classes, kind="yukawa")
        ref = oracle_fmm_h2d(tree, trav, q, order, 2j)
        e_or = rel(pot, ref)
        errs.append(float(np.max(np.abs(pot - direct))/np.max(np.abs(direct))))
        print(f"Y2D classes={use_classes} p={order}: vs oracle FMM {e_or:.2e}, "
              f"rel-inf vs direct {errs[-1]:.2e}")
        assert e_or < 1e-12
    assert errs[1] < errs[0] < 1e-6


@pytest.mark.parametrize("use_classes", [True, False])
def test_h2d_fmm_complex_k(actx, use_classes):
    """The 2D Helmholtz FMM with the reference's evanescent k = 0.2 (0.707 +
    0.707 i) against the oracle FMM (same tree and lists) and the direct
    sum."""
    from sumpy_amd.tree import build_tree_and_traversal
    k = 0.2*(0.707 + 0.707j)
    src, tgt, q = _fmm_case()
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    direct = oh.p2p(k, tgt, src, q)
    errs = []
    for order in (4, 8):
        pot = _run_fmm(actx, tree, trav, q, order, k, use_classes, kind="helmholtz_c")
        ref = oracle_fmm_h2d(tree, trav, q, order, k)
        e_or = rel(pot, ref)
        errs.append(float(np.max(np.abs(pot - direct))/np.max(np.abs(direct))))
        print(f"complex k classes={use_classes} p={order}: vs oracle FMM {e_or:.2e}, "
              f"rel-inf vs direct {errs[-1]:.2e}")
        assert e_or < 1e-12
    assert errs[1] < 0.1*errs[0] and errs[0] < 1e-3


def _der_kernels(kind):
    from sumpy_amd.kernel import AxisTargetDerivative, DirectionalSourceDerivative
    knl, mcls, lcls, kname, arg = _kernel(kind)
    skn = [knl, DirectionalSourceDerivative(knl, "dir_vec")]
    tkn = [knl, AxisTargetDerivative(0, knl), AxisTargetDerivative(1, knl)]
    return knl, mcls, lcls, kname, arg, skn, tkn


@pytest.mark.parametrize(("kind", "k"), [("helmholtz", 1.3), ("helmholtz_c", 0.2*(0.707 + 0.707j)),
                                         ("yukawa", 2.0)])
def test_h2d_p2p_dipoles_and_gradients_match_oracle(actx, kind, k):
    """P2P (dense and list-1 CSR) with charge + DirectionalSourceDerivative
    sources (strength_usage [0, 1]) and potential + d/dx + d/dy targets
    against the oracle."""
    from sumpy_amd import P2P, P2PFromCSR
    knl, _, _, kname, arg, skn, tkn = _der_kernels(kind)
    rng = np.random.default_rng(12)
    src, tgt = 4*rng.random((2, 500)), 4*rng.random((2, 200))
    q0 = rng.standard_normal(500) + 1j*rng.standard_normal(500)
    q1 = rng.standard_normal(500) + 1j*rng.standard_normal(500)
    n = rng.standard_normal((2, 500))
    ref = oh.p2p(arg(k), tgt, src, q0, dipoles=(q1*n[0], q1*n[1]), grad=True)
    d = actx.from_numpy
    print(f"h2d p2p derivatives {kind} k={k}", flush=True)
    got = P2P(tkn, exclude_self=False, source_kernels=skn, strength_usage=[0, 1])(
        actx, d(tgt), d(src), [d(q0), d(q1)], dir_vec=d(n), **{kname: k})
    for g, r in zip(got, ref):
        assert rel(actx.to_numpy(g), r) < 1e-12
    # CSR: one target box (all targets) with one source box (all sources)
    i32 = lambda a: d(np.asarray(a, np.int32))  # noqa: E731
    got = P2PFromCSR(tkn, exclude_self=False, source_kernels=skn, strength_usage=[0, 1])(
        actx, d(tgt), d(src), target_boxes=i32([0]), source_box_starts=i32([0, 1]),
        source_box_lists=i32([0]), box_source_starts=i32([0]),
        box_source_counts_nonchild=i32([500]), box_target_starts=i32([0]),
        box_target_counts_nonchild=i32([200]), strength=[d(q0), d(q1)], dir_vec=d(n),
        **{kname: k})
    for g, r in zip(got, ref):
        assert rel(actx.to_numpy(g), r) < 1e-12


@pytest.mark.parametrize(("kind", "order", "k"), [
    ("helmholtz", 10, 0.05), ("helmholtz_c", 12, 0.2*(0.707 + 0.707j)), ("yukawa", 10, 2.0)])
def test_h2d_dipole_p2e_and_gradient_e2p_match_oracle(actx, kind, order, k):
    """P2M and P2L with charge + dipole sources, M2P and L2P with potential +
    gradient targets, through the operators, against the oracle."""
    from sumpy_amd import E2PFromCSR, E2PFromSingleBox, P2EFromCSR, P2EFromSingleBox
    knl, mcls, lcls, kname, arg, skn, tkn = _der_kernels(kind)
    mpole, local = mcls(knl, order), lcls(knl, order)
    kw = {kname: k}
    ka = arg(k)
    src, _, tgt, c = _chain_geometry()
    rng = np.random.default_rng(3)
    q0 = rng.standard_normal(15) + 1j*rng.standard_normal(15)
    q1 = rng.standard_normal(15) + 1j*rng.standard_normal(15)
    n = rng.standard_normal((2, 15))
    dip = (q1*n[0], q1*n[1])
    d = actx.from_numpy
    i32 = lambda a: d(np.asarray(a, np.int32))  # noqa: E731
    centers = d(np.ascontiguousarray(c))
    bss, bsc = i32([0, 15, 15, 15]), i32([15, 0, 0, 0])
    bts, btc = i32([0, 0, 0, 0]), i32([0, 0, 0, 400])
    ncoef = 2*order + 1
    print(f"h2d dipole/gradient operators {kind} p={order}", flush=True)
    M = torch.zeros((4, ncoef), dtype=torch.complex128, device=actx.device)
    P2EFromSingleBox(mpole, kernels=skn, strength_usage=[0, 1])(
        actx, sources=d(src), centers=centers, strengths=[d(q0), d(q1)], dir_vec=d(n),
        tgt_expansions=M, tgt_base_ibox=0, rscale=0.5, source_boxes=i32([0]),
        box_source_starts=bss, box_source_counts_nonchild=bsc, **kw)
    m_ref = oh.p2m(order, ka, c[:, 0], src, q0, 0.5, dipoles=dip)
    assert rel(actx.to_numpy(M[0]), m_ref) < 1e-12
    L = torch.zeros((4, ncoef), dtype=torch.complex128, device=actx.device)
    P2EFromCSR(local, kernels=skn, strength_usage=[0, 1])(
        actx, sources=d(src), centers=centers, strengths=[d(q0), d(q1)], dir_vec=d(n),
        tgt_expansions=L, tgt_base_ibox=0, rscale=0.25, target_boxes=i32([3]),
        source_box_starts=i32([0, 1]), source_box_lists=i32([0]), box_source_starts=bss,
        box_source_counts_nonchild=bsc, **kw)
    l_ref = oh.p2l(order, ka, c[:, 3], src, q0, 0.25, dipoles=dip)
    assert rel(actx.to_numpy(L[3]), l_ref) < 1e-12
    got = E2PFromSingleBox(local, tkn)(
        actx, src_expansions=L, src_base_ibox=0, target_boxes=i32([3]), centers=centers,
        targets=d(tgt), box_target_starts=bts, box_target_counts_nonchild=btc, rscale=0.25,
        **kw)
    ref = oh.l2p(order, ka, c[:, 3], l_ref, tgt, 0.25, grad=True)
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
    ref = oh.m2p(order, ka, c[:, 0], m_ref, tgt, 0.5, grad=True)
    for g, r in zip(got, ref):
        assert rel(actx.to_numpy(g), r) < 1e-12
    # the expansions reproduce the direct sum at these distances
    direct = oh.p2p(ka, tgt, src, q0, dipoles=dip, grad=True)
    for g, r in zip(got, direct):
        assert rel(actx.to_numpy(g), r) < 1e-6


@pytest.mark.parametrize("kind", ["helmholtz", "yukawa"])
def test_h2d_fmm_dipoles_and_gradients(actx, kind):
    """The 2D Bessel FMM with charge + dipole sources and potential +
    gradient targets (the reference's test_sumpy_fmm setup) against the
    direct sum, converging with the order."""
    from functools import partial

    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.tree import build_tree_and_traversal
    knl, mcls, lcls, kname, arg, skn, tkn = _der_kernels(kind)
    k = 0.05 if kind == "helmholtz" else 2.0
    src, tgt, q = _fmm_case()
    rng = np.random.default_rng(21)
    q1 = rng.standard_normal(1000)
    n = rng.standard_normal((2, 1000))
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    ids = tree.host["user_source_ids"]
    ref = oh.p2p(arg(k), tgt, src, q, dipoles=(q1*n[0], q1*n[1]), grad=True)
    errs = []
    for order in (6, 10):
        ti = SumpyTreeIndependentDataForWrangler(actx, partial(mcls, knl), partial(lcls, knl),
                                                 tkn, source_kernels=skn, strength_usage=[0, 1])
        wr = SumpyExpansionWrangler(ti, trav, np.complex128, lambda *a: order,
                                    kernel_extra_kwargs={kname: k},
                                    source_extra_kwargs={"dir_vec": actx.from_numpy(
                                        np.ascontiguousarray(n[:, ids]))})
        print(f"h2d fmm derivatives {kind} p={order}", flush=True)
        pots = drive_fmm(actx, wr, [actx.from_numpy(q.astype(np.complex128)),
                                    actx.from_numpy(q1.astype(np.complex128))])
        errs.append(max(rel(actx.to_numpy(p), r) for p, r in zip(pots, ref)))
    print(f"{kind}: dipole/gradient FMM errors {errs}")
    assert errs[1] < 0.2*errs[0] and errs[0] < 1e-3


@pytest.mark.parametrize("use_classes", [True, False])
def test_h2d_fmm_order_varies_with_level(actx, use_classes):
    """The reference's H2D case with ``order_varies_with_level``
    (``sumpy/test/test_fmm.py:100-101,268-270``: order + lev % 2): M2M, L2L
    and M2L between different orders; converging to the direct sum and
    between the uniform-order results."""
    from functools import partial

    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.tree import build_tree_and_traversal
    knl, mcls, lcls, kname, _ = _kernel("helmholtz")
    src, tgt, q = _fmm_case()
    tree, trav = build_tree_and_traversal(actx, src, tgt, 30)
    direct = oh.p2p(0.05, tgt, src, q)
    errs = []
    for order in (4, 5):
        ti = SumpyTreeIndependentDataForWrangler(actx, partial(mcls, knl), partial(lcls, knl),
                                                 [knl])
        wr = SumpyExpansionWrangler(ti, trav, np.complex128,
                                    lambda kernel, args, tree, lev, o=order: o + lev % 2,
                                    kernel_extra_kwargs={kname: 0.05},
                                    _disable_translation_classes=not use_classes)
        print(f"h2d varying orders {order} + lev % 2, classes={use_classes}", flush=True)
        pot, = drive_fmm(actx, wr, [actx.from_numpy(q.astype(np.complex128))])
        errs.append(float(np.max(np.abs(actx.to_numpy(pot) - direct))/np.max(np.abs(direct))))
    uni = [float(np.max(np.abs(_run_fmm(actx, tree, trav, q, o, 0.05, use_classes) - direct))
                 / np.max(np.abs(direct))) for o in (4, 6)]
    print(f"varying-order errors {errs}, uniform orders 4 and 6 {uni}")
    assert errs[1] < errs[0] < 1e-3
    assert uni[1] <= errs[0] <= uni[0] and uni[1] <= errs[1] <= uni[0]
