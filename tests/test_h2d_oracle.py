"""Pins of the H2D (2D Helmholtz Hankel/Bessel) oracle against what the
mathematics fixes: every chain of expansions and translations converges to
the direct sum (i/4) sum_s q_s H0(k|t - s|) (Graf's addition theorem) as the
order grows, the rscale choices cancel, and the Toeplitz class data equals
the per-pair M2L.  Geometry of the reference's ``test_translations``
(``sumpy/test/test_kernels.py:493-640``): 15 sources in a 0.7-wide square
around (2, 1), a 20 x 20 target grid of width 0.3 around (7.5, 1), k = 0.05,
multipole centres at the source box and (-0.2, 0.1) from it, local centres at
(0.3, -0.2) from the target grid and at its centre, rscales 0.5 and 0.25.
A dropped (-1)^j, a conjugated phase or a swapped J/H fails these chains.
"""

import numpy as np
import pytest

from oracle import h2d

K = 0.05


def _setup(k=K):
    rng = np.random.default_rng(18)
    origin = np.array([2.0, 1.0])
    src = 0.7*(-0.5 + rng.random((2, 15))) + origin[:, None]
    q = np.ones(15)/15
    g = np.linspace(-0.15, 0.15, 20)
    off = np.array([5.5, 0.0])
    tx, ty = np.meshgrid(g + off[0] + origin[0], g + off[1] + origin[1], indexing="ij")
    tgt = np.array([tx.ravel(), ty.ravel()])
    centers = np.array([[0, 0], [-0.2, 0.1], off + [0.3, -0.2], off]).T + origin[:, None]
    return src, q, tgt, centers, h2d.p2p(k, tgt, src, q)


def _chains(order, k=K, rs=(0.5, 0.25, 0.5, 0.25)):
    src, q, tgt, c, ref = _setup(k)
    m1r, m2r, l1r, l2r = rs
    m1 = h2d.p2m(order, k, c[:, 0], src, q, m1r)
    m2 = h2d.m2m(order, order, k, m1, c[:, 0], c[:, 1], m1r, m2r)
    l1 = h2d.m2l(order, order, k, m2, c[:, 2] - c[:, 1], m2r, l1r)
    l2 = h2d.l2l(order, order, k, l1, c[:, 2], c[:, 3], l1r, l2r)
    out = {"p2m2p": h2d.m2p(order, k, c[:, 0], m1, tgt, m1r),
           "p2m2m2p": h2d.m2p(order, k, c[:, 1], m2, tgt, m2r),
           "p2m2m2l2p": h2d.l2p(order, k, c[:, 2], l1, tgt, l1r),
           "full": h2d.l2p(order, k, c[:, 3], l2, tgt, l2r),
           "p2l2p": h2d.l2p(order, k, c[:, 3], h2d.p2l(order, k, c[:, 3], src, q, l2r),
                            tgt, l2r)}
    return {n: np.linalg.norm(v - ref)/np.linalg.norm(ref) for n, v in out.items()}, ref


def test_h2d_chains_converge_to_direct_sum():
    errs = [_chains(p)[0] for p in (2, 3, 4, 6, 10)]
    for name in errs[0]:
        e = [x[name] for x in errs]
        assert all(b < 0.5*a for a, b in zip(e, e[1:])), (name, e)
        assert e[-1] < 1e-12, (name, e)


@pytest.mark.parametrize("k", [1.3, 7.0])
def test_h2d_chains_at_larger_wavenumbers(k):
    """k |d| up to about 40 at the M2L (the Hankel functions of order
    2p there vary over many decades)."""
    errs = _chains(24, k)[0]
    assert max(errs.values()) < 1e-9, errs


def test_h2d_rscale_cancels():
    a, _ = _chains(8)
    b, _ = _chains(8, rs=(1.0, 0.1, 0.03, 2.0))
    for name in a:
        assert abs(a[name] - b[name]) <= 1e-3*a[name] + 1e-15, name


def test_h2d_class_data_is_toeplitz_of_per_pair_m2l():
    rng = np.random.default_rng(3)
    p = 7
    M = rng.standard_normal(2*p + 1) + 1j*rng.standard_normal(2*p + 1)
    d = np.array([0.9, -2.1])
    data = h2d.m2l_class_data(p, p, 0.8, d)
    assert data.shape == (4*p + 1,)
    # data[n + 2p] = H_n e^{i n theta}, H_{-n} = (-1)^n H_n
    n = np.arange(1, 2*p + 1)
    th = np.arctan2(d[1], d[0])
    assert np.allclose(data[2*p - n]*np.exp(1j*n*th), (-1.0)**n*data[2*p + n]*np.exp(-1j*n*th),
                       rtol=1e-13, atol=0)
    a = h2d.m2l(p, p, 0.8, M, d, 0.3, 0.2)
    b = h2d.m2l(p, p, 0.8, M, d, 0.3, 0.2, data=data)
    assert np.array_equal(a, b)


def test_h2d_p2p_matches_hankel_definition():
    """H0^(1) = J0 + i Y0 (scipy) and the i/4 scaling on one pair."""
    from scipy import special
    v = h2d.p2p(2.0, np.array([[0.3], [0.4]]), np.zeros((2, 1)), np.array([1.0]))[0]
    assert abs(v - 0.25j*(special.j0(1.0) + 1j*special.y0(1.0))) < 1e-15


def test_y2d_p2p_is_k0_over_2pi():
    """(i/4) H0(i lam r) = K0(lam r)/(2 pi) for the Yukawa kernel
    (``sumpy/kernel.py:879-884``)."""
    from scipy import special
    r = np.array([0.01, 0.3, 1.0, 7.5])
    v = np.array([h2d.p2p(2j, np.array([[x], [0.0]]), np.zeros((2, 1)), np.array([1.0]))[0]
                  for x in r])
    assert np.allclose(v, special.k0(2*r)/(2*np.pi), rtol=1e-14, atol=0)


@pytest.mark.parametrize("lam", [0.05, 2.0])
def test_y2d_chains_converge_to_direct_sum(lam):
    """The Yukawa expansions (Bessel argument i lam) through the same chains;
    the M2L's Hankel row grows like K_n, so the rscales keep the terms in
    range."""
    errs = [_chains(p, 1j*lam)[0] for p in (2, 4, 8, 14)]
    for name in errs[0]:
        e = [x[name] for x in errs]
        assert all(b < 0.5*a for a, b in zip(e, e[1:]) if a > 1e-13), (name, e)
        assert e[-1] < 1e-11, (name, e)


@pytest.mark.parametrize("k", [0.2*(0.707 + 0.707j), 1.3 + 0.4j])
def test_h2d_chains_converge_complex_k(k):
    """Damped waves (complex k; the first value is the reference's evanescent
    k, ``sumpy/test/test_kernels.py:168``): Graf's addition theorem holds for
    complex arguments, so the chains still converge to the direct sum."""
    errs = [_chains(p, k)[0] for p in (2, 4, 8, 16)]
    for name in errs[0]:
        e = [x[name] for x in errs]
        assert all(b < 0.5*a for a, b in zip(e, e[1:])), (name, e)
        assert e[-1] < 1e-10, (name, e)


@pytest.mark.parametrize("k", [1.3, 0.2*(0.707 + 0.707j), 2.0j])
def test_h2d_dipoles_and_gradients_are_derivatives(k):
    """The dipole (n . grad_source) and target-gradient forms of P2P, P2M,
    P2L, M2P and L2P equal central differences (h = 1e-5) of the charge /
    potential forms (k real, complex and the Yukawa scale 2i)."""
    rng = np.random.default_rng(4)
    src = rng.random((2, 6))
    tgt = np.array([[3.1, 2.8, 3.4], [0.2, -0.3, 0.5]])
    q = rng.standard_normal(6) + 1j*rng.standard_normal(6)
    n = rng.standard_normal((2, 6))
    n /= np.linalg.norm(n, axis=0)
    h = 1e-5
    c0, c1 = np.array([0.5, 0.5]), np.array([3.2, 0.1])

    def fd_src(f):
        return (f(src + h*n) - f(src - h*n))/(2*h)

    dip = (q*n[0], q*n[1])
    zero = np.zeros(6)
    # P2P: the dipole potential, and the gradients
    pot, gx, gy = h2d.p2p(k, tgt, src, zero, dipoles=dip, grad=True)
    ref = fd_src(lambda s: h2d.p2p(k, tgt, s, q))
    assert np.allclose(pot, ref, rtol=1e-7, atol=1e-9)
    for ax, g in enumerate((gx, gy)):
        e = np.eye(2)[:, ax:ax + 1]
        ref = (h2d.p2p(k, tgt + h*e, src, zero, dipoles=dip)
               - h2d.p2p(k, tgt - h*e, src, zero, dipoles=dip))/(2*h)
        assert np.allclose(g, ref, rtol=1e-6, atol=1e-8), ax
    # P2M / P2L dipoles
    for fn, c in ((h2d.p2m, c0), (h2d.p2l, c1)):
        got = fn(6, k, c, src, zero, 0.7, dipoles=dip)
        ref = fd_src(lambda s: fn(6, k, c, s, q, 0.7))
        assert np.allclose(got, ref, rtol=1e-7, atol=1e-9), fn.__name__
    # M2P / L2P gradients
    coeffs = rng.standard_normal(13) + 1j*rng.standard_normal(13)
    for fn, c, t in ((h2d.m2p, c0, tgt), (h2d.l2p, c1, tgt)):
        pot, gx, gy = fn(6, k, c, coeffs, t, 0.7, grad=True)
        assert np.allclose(pot, fn(6, k, c, coeffs, t, 0.7), rtol=1e-14)
        for ax, g in enumerate((gx, gy)):
            e = np.eye(2)[:, ax:ax + 1]
            ref = (fn(6, k, c, coeffs, t + h*e, 0.7) - fn(6, k, c, coeffs, t - h*e, 0.7))/(2*h)
            assert np.allclose(g, ref, rtol=1e-7, atol=1e-9), (fn.__name__, ax)


@pytest.mark.parametrize("k", [0.05, 0.2*(0.707 + 0.707j)])
def test_h2d_dipole_gradient_chain_converges(k):
    """Dipole sources through P2M -> M2M -> M2L -> L2L -> L2P with target
    gradients converge to the direct dipole-gradient sum (the reference's
    ``test_p2e2p`` with source derivative and the x derivative)."""
    src, q, tgt, c, _ = _setup(k)
    rng = np.random.default_rng(9)
    n = rng.standard_normal((2, src.shape[1]))
    n /= np.linalg.norm(n, axis=0)
    dip = (q*n[0], q*n[1])
    zero = np.zeros_like(q)
    ref = h2d.p2p(k, tgt, src, zero, dipoles=dip, grad=True)
    errs = []
    for order in (4, 8, 12):
        m1 = h2d.p2m(order, k, c[:, 0], src, zero, 0.5, dipoles=dip)
        m2 = h2d.m2m(order, order, k, m1, c[:, 0], c[:, 1], 0.5, 0.25)
        l1 = h2d.m2l(order, order, k, m2, c[:, 2] - c[:, 1], 0.25, 0.5)
        l2 = h2d.l2l(order, order, k, l1, c[:, 2], c[:, 3], 0.5, 0.25)
        got = h2d.l2p(order, k, c[:, 3], l2, tgt, 0.25, grad=True)
        errs.append(max(np.linalg.norm(g - r)/np.linalg.norm(r) for g, r in zip(got, ref)))
    assert errs[2] < 1e-3*errs[0] and errs[2] < 1e-9, errs
