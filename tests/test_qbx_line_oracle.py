"""Pins of the line-Taylor layer-potential oracle (``oracle/qbx_line.py``)
for the Laplace, Helmholtz and Yukawa kernels in 2D and 3D.

* 2D Laplace equals the independent C restatement ``oracle.qbx2d`` (Taylor
  jets of log q) for the single and double layer, with and without the
  target derivative;
* 3D Laplace equals the Legendre closed form
  1/|a + tau b| = sum_i tau^i (-1)^i |b|^i P_i(a.b/|a||b|) / |a|^{i+1};
* every kernel and every source/target derivative combination converges
  geometrically in the order to the kernel at the target, D_s K(a + b), when
  |b| < |a| (the line Taylor series at tau = 1 inside its radius |a|/|b|);
* the double layer is minus the directional derivative in the source
  position (finite differences of the single-layer value).

A dropped derivative term, a wrong sign or a swapped a/b fails one of these.
"""

import numpy as np
import pytest

import oracle
from oracle import qbx_line as ql

K = {"lap2": 0.0, "lap3": 0.0, "helm2": 1.3, "helm3": 1.3, "yuk2": 0.9, "yuk3": 0.9}


def _pair(dim, ratio, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(dim)
    a /= np.linalg.norm(a)
    b = rng.standard_normal(dim)
    b *= ratio/np.linalg.norm(b)
    n = rng.standard_normal(dim)
    n /= np.linalg.norm(n)
    return a, b, n


def test_lap2_matches_c_oracle():
    rng = np.random.default_rng(5)
    nt, ns, order = 3, 4, 5
    src = rng.standard_normal((2, ns))
    ctr = rng.standard_normal((2, nt))*0.2 + 3.0
    tgt = ctr + 0.3*rng.standard_normal((2, nt))
    nrm = rng.standard_normal((2, ns))
    nrm /= np.linalg.norm(nrm, axis=0)
    dens = rng.standard_normal(ns)
    for srcmode in (0, 1):
        for ax in (-1, 0, 1):
            ref = oracle.qbx2d(order, srcmode, ax, tgt, ctr, np.ones(nt), src,
                               nrm if srcmode else None, dens)
            got = ql.layer_potential("lap2", 0.0, order, tgt, ctr, src, dens,
                                     normals=nrm if srcmode else None, tgt_axis=ax)
            assert np.allclose(got.real, ref, rtol=1e-12, atol=1e-14), (srcmode, ax)
            assert np.all(got.imag == 0)


def test_lap3_matches_legendre_closed_form():
    from scipy.special import eval_legendre
    a, b, _ = _pair(3, 0.6, 1)
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    u = a @ b/(na*nb)
    for order in (0, 1, 4, 9):
        ref = sum((-1)**i*nb**i*eval_legendre(i, u)/na**(i + 1) for i in range(order + 1))
        got = ql.pair_value("lap3", 0.0, order, a, b)
        assert abs(got - ref) <= 1e-13*abs(ref), order


@pytest.mark.parametrize("kind", ql.KINDS)
@pytest.mark.parametrize("mode", ["slp", "dlp", "slp_dx", "dlp_dy"])
def test_converges_to_kernel_at_target(kind, mode):
    dim = 3 if kind.endswith("3") else 2
    a, b, n = _pair(dim, 0.15, 7)
    nrm = n if mode.startswith("dlp") else None
    ax = {"slp": -1, "dlp": -1, "slp_dx": 0, "dlp_dy": 1}[mode]
    ref = ql.direct_value(kind, K[kind], a + b, nrm, ax)
    errs = [abs(ql.pair_value(kind, K[kind], p, a, b, nrm, ax) - ref)/abs(ref)
            for p in (2, 5, 9)]
    # geometric in the order (a term can vanish for one direction, so
    # compare the lowest and the highest order only)
    assert errs[-1] < 1e-3*errs[0] and errs[-1] < 1e-6, errs


@pytest.mark.parametrize("kind", ["lap3", "helm2", "helm3", "yuk2", "yuk3"])
def test_dlp_is_minus_source_directional_derivative(kind):
    """D_s = -n.grad_a of the single-layer coefficients, a = c - s: moving the
    source by +h n changes a by -h n."""
    dim = 3 if kind.endswith("3") else 2
    a, b, n = _pair(dim, 0.3, 11)
    h = 1e-5
    order = 4
    fd = (ql.pair_value(kind, K[kind], order, a - h*n, b)
          - ql.pair_value(kind, K[kind], order, a + h*n, b))/(2*h)
    dlp = ql.pair_value(kind, K[kind], order, a, b, n)
    # d/ds_n of the value = -d/da_n; the DLP kernel is n . grad_s = -n . grad_a
    assert abs(dlp - fd) <= 1e-8*max(1.0, abs(dlp)), (dlp, fd)


def test_yuk2_is_k0_over_2pi_and_helm2_is_hankel():
    from scipy import special
    d = np.array([0.4, 0.3])
    v = ql.SCALE["yuk2"]*ql.direct_value("yuk2", 2.0, d)
    assert abs(v - special.k0(2.0*0.5)/(2*np.pi)) < 1e-14
    v = ql.SCALE["helm2"]*ql.direct_value("helm2", 2.0, d)
    assert abs(v - 0.25j*special.hankel1(0, 1.0)) < 1e-14


@pytest.mark.parametrize("mode", ["slp", "dlp_dy"])
def test_helm2_complex_k_converges_to_kernel_at_target(mode):
    """2D Helmholtz with a complex (damped) k: the line expansion still
    converges to the kernel at the target."""
    k = 1.3 + 0.4j
    a, b, n = _pair(2, 0.15, 7)
    nrm = n if mode.startswith("dlp") else None
    ax = -1 if mode == "slp" else 1
    ref = ql.direct_value("helm2", k, a + b, nrm, ax)
    errs = [abs(ql.pair_value("helm2", k, p, a, b, nrm, ax) - ref)/abs(ref) for p in (2, 5, 9)]
    assert errs[-1] < 1e-3*errs[0] and errs[-1] < 1e-6, errs
