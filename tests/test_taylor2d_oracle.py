"""Pins of the 2D volume-Taylor oracle (``oracle/taylor2d.py``) against what
the mathematics fixes, for the Laplace and Helmholtz kernels, full and
PDE-compressed:

* the derivatives equal the closed forms d/dx log r = x/r^2 and
  d/dx H0(k r) = -k H1(k r) x/r;
* every chain of expansions and translations converges to the direct sum as
  the order grows (the geometry of the reference's ``test_translations``,
  ``sumpy/test/test_kernels.py:493-640``: 15 sources in a 0.7-wide square
  around (2, 1), a target grid of width 0.3 around (7.5, 1));
* compressed and full expansions give the same potentials (the PDE
  projection reproduces every derivative of a PDE solution);
* the rscale choices cancel.
A dropped projection term, a wrong sign of a = c - s or a transposed
re-expansion fails one of these.
"""

import numpy as np
import pytest

from oracle import taylor2d as t2

K = 0.05
KC = 0.2*(0.707 + 0.707j)   # the reference's evanescent k (test_kernels.py:168)


def test_derivatives_match_closed_forms():
    from scipy.special import hankel1
    d = np.array([0.7, -0.4])
    r = np.hypot(*d)
    assert abs(t2.deriv("lap", 0, d, 1, 0) - d[0]/r**2) < 1e-15
    assert abs(t2.deriv("lap", 0, d, 0, 1) - d[1]/r**2) < 1e-15
    v = t2.deriv("helm", 2.0, d, 1, 0)
    assert abs(v - (-2.0*hankel1(1, 2.0*r)*d[0]/r)) < 1e-14
    # the PDEs: Laplacian log r = 0, Laplacian H0(k r) = -k^2 H0(k r)
    assert abs(t2.deriv("lap", 0, d, 2, 0) + t2.deriv("lap", 0, d, 0, 2)) < 1e-13
    lap = t2.deriv("helm", 2.0, d, 2, 0) + t2.deriv("helm", 2.0, d, 0, 2)
    assert abs(lap + 4.0*t2.deriv("helm", 2.0, d, 0, 0)) < 1e-13


@pytest.mark.parametrize(("kind", "nmax", "k"), [("lap", 8, 0), ("helm", 4, 1.3),
                                              ("helm", 3, 1.3 + 0.4j)])
def test_derivatives_match_sympy(kind, nmax, k):
    """Exact differentiation of log r / H0(k r) (sympy) at every order up to
    nmax, for a real and a complex (damped) k."""
    import sympy as sp
    x, y = sp.symbols("x y", real=True)
    r = sp.sqrt(x**2 + y**2)
    ks = sp.Rational(13, 10) + (sp.Rational(2, 5)*sp.I if isinstance(k, complex) else 0)
    g = sp.log(r) if kind == "lap" else sp.hankel1(0, ks*r)
    d = np.array([0.7, -0.4])
    for n in range(nmax + 1):
        for my in range(n + 1):
            e = sp.diff(g, x, n - my, y, my) if n else g
            ref = complex(sp.N(e.subs({x: sp.Rational(7, 10), y: sp.Rational(-4, 10)}), 20))
            got = t2.deriv(kind, k, d, n - my, my)
            assert abs(got - ref) <= 1e-12*max(1.0, abs(ref)), (n, my, got, ref)


def _setup(kind, k=K):
    rng = np.random.default_rng(18)
    origin = np.array([2.0, 1.0])
    src = 0.7*(-0.5 + rng.random((2, 15))) + origin[:, None]
    q = np.ones(15)/15
    g = np.linspace(-0.15, 0.15, 5)
    off = np.array([5.5, 0.0])
    tx, ty = np.meshgrid(g + off[0] + origin[0], g + off[1] + origin[1], indexing="ij")
    tgt = np.array([tx.ravel(), ty.ravel()])
    c = np.array([[0, 0], [-0.2, 0.1], off + [0.3, -0.2], off]).T + origin[:, None]
    return src, q, tgt, c, t2.p2p(kind, k, tgt, src, q)


def _chains(kind, compressed, p, rs=(0.5, 0.25, 0.5, 0.25), K=K):
    src, q, tgt, c, ref = _setup(kind, K)
    m1r, m2r, l1r, l2r = rs
    m1 = t2.p2m(p, compressed, kind, K, c[:, 0], src, q, m1r)
    m2 = t2.m2m(p, p, compressed, kind, K, m1, c[:, 0], c[:, 1], m1r, m2r)
    # (the oracle's M2L takes one rscale for both sides)
    l1 = t2.m2l(p, p, compressed, kind, K, m2, c[:, 2] - c[:, 1], m2r) if m2r == l1r else None
    out = {"p2m2p": t2.m2p(p, compressed, kind, K, c[:, 0], m1, tgt, m1r),
           "p2m2m2p": t2.m2p(p, compressed, kind, K, c[:, 1], m2, tgt, m2r),
           "p2l2p": t2.l2p(p, compressed, kind, K, c[:, 3],
                           t2.p2l(p, compressed, kind, K, c[:, 3], src, q, l2r), tgt, l2r)}
    if l1 is not None:
        l2 = t2.l2l(p, p, compressed, kind, K, l1, c[:, 2], c[:, 3], l1r, l2r)
        out["p2m2m2l2p"] = t2.l2p(p, compressed, kind, K, c[:, 2], l1, tgt, l1r)
        out["full"] = t2.l2p(p, compressed, kind, K, c[:, 3], l2, tgt, l2r)
    return {n: np.linalg.norm(v - ref)/np.linalg.norm(ref) for n, v in out.items()}


@pytest.mark.parametrize("kind", ["lap", "helm"])
@pytest.mark.parametrize("compressed", [True, False])
def test_chains_converge_to_direct_sum(kind, compressed):
    errs = [_chains(kind, compressed, p, rs=(0.5, 0.25, 0.25, 0.25)) for p in (1, 3, 5)]
    for name in errs[0]:
        e = [x[name] for x in errs]
        assert all(b < 0.2*a for a, b in zip(e, e[1:])), (name, e)
        assert e[-1] < 1e-5, (name, e)


@pytest.mark.parametrize("kind", ["lap", "helm"])
def test_compressed_equals_full(kind):
    src, q, tgt, c, _ = _setup(kind)
    p = 4
    v = [t2.l2p(p, cm, kind, K, c[:, 3], t2.p2l(p, cm, kind, K, c[:, 3], src, q, 0.3), tgt, 0.3)
         for cm in (True, False)]
    assert np.allclose(v[0], v[1], rtol=1e-12, atol=0)
    v = [t2.m2p(p, cm, kind, K, c[:, 0], t2.p2m(p, cm, kind, K, c[:, 0], src, q, 0.3), tgt, 0.3)
         for cm in (True, False)]
    assert np.allclose(v[0], v[1], rtol=1e-12, atol=0)


@pytest.mark.parametrize("kind", ["lap", "helm"])
def test_rscale_cancels(kind):
    a = _chains(kind, True, 3, rs=(0.5, 0.25, 0.25, 0.25))
    b = _chains(kind, True, 3, rs=(1.0, 0.1, 0.1, 2.0))
    for name in a:
        assert abs(a[name] - b[name]) <= 1e-6*a[name] + 1e-14, name


@pytest.mark.parametrize("compressed", [True, False])
def test_chains_converge_complex_k(compressed):
    """Damped waves (complex k, the reference's evanescent value): the same
    convergence, and the compressed projection with (k rho)^2 complex."""
    errs = [_chains("helm", compressed, p, rs=(0.5, 0.25, 0.25, 0.25), K=KC) for p in (1, 3, 5)]
    for name in errs[0]:
        e = [x[name] for x in errs]
        assert all(b < 0.2*a for a, b in zip(e, e[1:])), (name, e)
        assert e[-1] < 1e-5, (name, e)


@pytest.mark.parametrize(("kind", "k"), [("lap", 0.0), ("helm", 1.3), ("helm", KC)])
@pytest.mark.parametrize("compressed", [True, False])
def test_dipoles_and_gradients_are_derivatives(kind, k, compressed):
    """The dipole (n . grad_source) forms of P2P, P2M and P2L and the
    target-gradient forms of P2P, M2P and L2P equal central differences
    (h = 1e-5) of the charge / potential forms."""
    rng = np.random.default_rng(6)
    src = 0.3*rng.random((2, 5))
    tgt = np.array([[3.1, 2.8, 3.4], [0.2, -0.3, 0.5]])
    q = rng.standard_normal(5) + 0j
    n = rng.standard_normal((2, 5))
    dip = (q*n[0], q*n[1])
    zero = np.zeros(5, complex)
    h = 1e-5
    c0, c1 = np.array([0.15, 0.15]), np.array([3.1, 0.1])

    def fd_src(f):
        return (f(src + h*n) - f(src - h*n))/(2*h)

    pot, gx, gy = t2.p2p(kind, k, tgt, src, zero, dipoles=dip, grad=True)
    assert np.allclose(pot, fd_src(lambda s: t2.p2p(kind, k, tgt, s, q)), rtol=1e-7, atol=1e-9)
    for ax, g in enumerate((gx, gy)):
        e = np.eye(2)[:, ax:ax + 1]
        ref = (t2.p2p(kind, k, tgt + h*e, src, zero, dipoles=dip)
               - t2.p2p(kind, k, tgt - h*e, src, zero, dipoles=dip))/(2*h)
        assert np.allclose(g, ref, rtol=1e-6, atol=1e-8), ax
    for fn, c in ((t2.p2m, c0), (t2.p2l, c1)):
        got = fn(4, compressed, kind, k, c, src, zero, 0.5, dipoles=dip)
        ref = fd_src(lambda s: fn(4, compressed, kind, k, c, s, q, 0.5))
        assert np.allclose(got, ref, rtol=1e-7, atol=1e-9), fn.__name__
    coeffs = rng.standard_normal(len(t2.stored_ids(4, compressed, kind))) + 0j
    for fn, c in ((t2.m2p, c0), (t2.l2p, c1)):
        pot, gx, gy = fn(4, compressed, kind, k, c, coeffs, tgt, 0.5, grad=True)
        assert np.allclose(pot, fn(4, compressed, kind, k, c, coeffs, tgt, 0.5), rtol=1e-14)
        for ax, g in enumerate((gx, gy)):
            e = np.eye(2)[:, ax:ax + 1]
            ref = (fn(4, compressed, kind, k, c, coeffs, tgt + h*e, 0.5)
                   - fn(4, compressed, kind, k, c, coeffs, tgt - h*e, 0.5))/(2*h)
            assert np.allclose(g, ref, rtol=1e-6, atol=1e-8), (fn.__name__, ax)
