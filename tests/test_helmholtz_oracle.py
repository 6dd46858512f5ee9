"""Pins of the Helmholtz-3D oracle (CPU only).

* the scaled derivatives of ``e^{ik|d|}/|d|`` (the reference's radial
  recursion, ``sumpy/derivative_taker.py:290-382``) against sympy's exact
  differentiation at a rational point, every multi-index up to order 5;
* the Laplace limit ``k = 0`` against the oracle's Laplace derivatives;
* the PDE identity ``D_(m + 2e_x) + D_(m + 2e_y) + D_(m + 2e_z) = -(k rho)^2
  D_m`` at order 10 (a dropped or mis-scaled term fails it);
* the product's Helmholtz projection table against the oracle's independent
  restatement of ``get_projection_matrix`` (``sumpy/expansion/__init__.py:
  757-886``).
"""

import numpy as np
import pytest

import oracle
from oracle import helmholtz as oh
from oracle import tables


def test_derivs_against_sympy():
    sp = pytest.importorskip("sympy")
    x, y, z = sp.symbols("x y z", real=True)
    k = sp.Rational(7, 3)
    r = sp.sqrt(x**2 + y**2 + z**2)
    G = sp.exp(sp.I*k*r)/r
    pt = {x: sp.Rational(3, 5), y: sp.Rational(-4, 7), z: sp.Rational(5, 4)}
    rho = 0.37
    D = oh.derivs(np.array([float(v) for v in pt.values()]), rho, float(k), 5)
    worst = 0.0
    for mi, val in D.items():
        ex = sp.diff(G, x, mi[0], y, mi[1], z, mi[2]).subs(pt)
        ref = complex(sp.N(ex, 30))*rho**sum(mi)
        worst = max(worst, abs(val[0] - ref)/abs(ref))
    assert worst < 1e-12


def test_laplace_limit():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-2, 2, (3, 5)) + np.array([[3.0], [0], [0]])
    rho = 0.8
    D = oh.derivs(pts, rho, 0.0, 8)
    for j in range(pts.shape[1]):
        L = oracle.lap3d_derivs(pts[:, j], rho, 8, (8, 8, 8))
        for mi, v in D.items():
            assert abs(v[j].imag) < 1e-300 + 1e-14*abs(v[j].real)
            assert v[j].real == pytest.approx(L[mi], rel=1e-11, abs=1e-14*abs(L[0, 0, 0]))


def test_pde_identity():
    rng = np.random.default_rng(5)
    pts = rng.uniform(1, 3, (3, 4))
    rho, k = 0.3, 4.5
    D = oh.derivs(pts, rho, k, 12)
    for mi, v in D.items():
        if sum(mi) > 10:
            continue
        lap = sum(D[tuple(m + 2*(a == i) for a, m in enumerate(mi))] for i in range(3))
        assert np.allclose(lap, -(k*rho)**2*v, rtol=1e-9, atol=1e-12*np.abs(v).max())


@pytest.mark.parametrize("order", [2, 5, 12])
def test_projection_with_k(order):
    from sumpy_amd.expansion.terms import get_terms_table
    rho, k = 0.21, 10.0
    P = get_terms_table(order, 3, True).projection((k*rho)**2)
    Q = tables.wrangler(order, 3, True, helmholtz=True).projection(rho, k)
    assert np.allclose(P, Q, rtol=1e-13, atol=1e-13)
    assert not np.allclose(P, get_terms_table(order, 3, True).projection(0.0))


def _sphere(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((3, n))
    x /= np.linalg.norm(x, axis=0)
    return x*(1 + 1e-3*rng.standard_normal(n)), (rng.standard_normal(n)
                                                 + 1j*rng.standard_normal(n))


@pytest.mark.parametrize("case", ["uniform", "sphere"])
def test_oracle_helmholtz_fmm_converges(case):
    """The oracle-driven Helmholtz FMM against the direct sum: error falls
    with the order (truncation, the only approximation), down to ~1e-5 at
    order 8 for k * box size <= 1 (1.2e-5 uniform).  A sign, index or projection mistake in
    any stage leaves an O(1) error."""
    from oracle_fmm import oracle_fmm
    from sumpy_amd.array_context import TorchArrayContext
    from sumpy_amd.tree import build_tree_and_traversal
    if case == "uniform":
        rng = np.random.default_rng(1)
        src = rng.random((3, 1500))
        q = rng.standard_normal(1500) + 1j*rng.standard_normal(1500)
        k, mp = 2.0, 30
    else:
        src, q = _sphere(1500, 2)
        k, mp = 1.5, 30
    tree, trav = build_tree_and_traversal(TorchArrayContext("cpu"), src, None, mp)
    ref, _ = oracle.p2p_dense(13, src, src, q, t2s=np.arange(src.shape[1]), k=k)
    errs = []
    for order in (3, 8):
        pot = oracle_fmm(tree, trav, q, order, exclude_self=True, k=k)
        errs.append(np.linalg.norm(pot - ref)/np.linalg.norm(ref))
    assert errs[1] < 5e-5 and errs[1] < errs[0]/20, errs


@pytest.mark.parametrize("use_dipoles", [False, True])
def test_p2l_then_l2p_converges_to_direct(use_dipoles):
    """Oracle P2L (charges and dipoles) followed by L2P on the real and
    imaginary planes against the direct Helmholtz sum: the error falls
    geometrically with the order (pins the P2L dipole sign and scaling and the
    (k rho)^2 projection used by L2P)."""
    rng = np.random.default_rng(8)
    k = 3.0
    ns, nt = 40, 30
    src = rng.uniform(-0.5, 0.5, (3, ns)) + np.array([[3.0], [0.5], [0.0]])
    tgt = rng.uniform(-0.5, 0.5, (3, nt))
    q = rng.standard_normal(ns) + 1j*rng.standard_normal(ns)
    wd = ([rng.standard_normal(ns) + 1j*rng.standard_normal(ns) for _ in range(3)]
          if use_dipoles else None)
    ref, _ = oracle.p2p_dense(13, tgt, src, q, wd=wd, k=k)
    # one target box (id 0) with center 0 and every target; one source box (id 1)
    tree = {"box_source_starts": np.array([0, 0], np.int32),
            "box_source_counts_nonchild": np.array([0, ns], np.int32),
            "box_target_starts": np.array([0, 0], np.int32),
            "box_target_counts_nonchild": np.array([nt, 0], np.int32)}
    centers = np.zeros((3, 2))
    errs = []
    for order in (4, 8, 12):
        rho = 1.0/order
        n = (order + 1)**2
        L = np.zeros((1, n), np.complex128)
        oh.p2l(order, np.array([0]), np.array([0, 1]), np.array([1]), tree, centers, src,
               q, wd, rho, k, L, 0)
        parts = []
        for part in (L.real, L.imag):
            pot = np.zeros(nt)
            oracle.l2p(order, np.array([0], np.int32), tree, centers, tgt, rho,
                       np.ascontiguousarray(part), 0, pot, k=k)
            parts.append(pot)
        pot = parts[0] + 1j*parts[1]
        errs.append(np.linalg.norm(pot - ref)/np.linalg.norm(ref))
    assert errs[2] < 1e-5 and errs[2] < errs[1] < errs[0], errs


def test_p2m_then_m2p_gradient_converges_to_direct():
    """Oracle P2M (real and imaginary planes, (k rho)^2 projection) followed
    by M2P with target gradients against the direct Helmholtz potential and
    gradient of a far cluster."""
    rng = np.random.default_rng(9)
    k = 2.5
    ns, nt = 50, 20
    src = rng.uniform(-0.5, 0.5, (3, ns))
    tgt = rng.uniform(-0.5, 0.5, (3, nt)) + np.array([[0.0], [3.2], [0.4]])
    q = rng.standard_normal(ns) + 1j*rng.standard_normal(ns)
    ref, gref = oracle.p2p_dense(13, tgt, src, q, k=k, want_grad=True)
    tree = {"box_source_starts": np.array([0, 0], np.int32),
            "box_source_counts_nonchild": np.array([ns, 0], np.int32),
            "box_target_starts": np.array([0, 0], np.int32),
            "box_target_counts_nonchild": np.array([0, nt], np.int32)}
    centers = np.zeros((3, 2))
    centers[:, 1] = [0.0, 3.2, 0.4]
    errs = []
    for order in (4, 8, 12):
        rho = 1.0/order
        n = (order + 1)**2
        M = np.zeros((1, n), np.complex128)
        for part, w in (("real", q.real), ("imag", q.imag)):
            out = np.zeros((1, n))
            oracle.p2m(order, np.array([0], np.int32), tree, centers, src,
                       np.ascontiguousarray(w), None, rho, out, 0, k=k)
            M += out if part == "real" else 1j*out
        pot = np.zeros(nt, np.complex128)
        grad = [np.zeros(nt, np.complex128) for _ in range(3)]
        oh.m2p(order, np.array([1]), np.array([0, 1]), np.array([0]), tree, centers, tgt,
               rho, k, M, 0, pot, grad)
        e = [np.linalg.norm(pot - ref)/np.linalg.norm(ref)]
        e += [np.linalg.norm(g - gr)/np.linalg.norm(gr) for g, gr in zip(grad, gref)]
        errs.append(max(e))
    assert errs[2] < 1e-5 and errs[2] < errs[1] < errs[0], errs
