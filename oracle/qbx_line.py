"""CPU oracle for the line-Taylor layer potential of the Laplace, Helmholtz
and Yukawa kernels in 2D and 3D -- TEST INFRASTRUCTURE ONLY (imported by
``tests/`` alone).

Restates ``LineTaylorLocalExpansion.coefficients_from_source`` /
``evaluate`` (``sumpy/expansion/local.py:132-185``) as used by
``LayerPotential`` (``sumpy/qbx.py:109-147,262-308``): per (target, source)
the coefficients are ``D_s d^i/dtau^i K(a + tau b)`` at ``tau = 0``, with
``a = c - s`` and ``b = t - c`` (``qbx.py:284-285``), and the value is
``sum_{i<=p} coeff_i / i!``.  ``D_s`` is the source derivative
(``-n . grad_a`` for ``DirectionalSourceDerivative``, ``kernel.py:2474-2490``;
``-d/da_j`` for ``AxisSourceDerivative``); an ``AxisTargetDerivative``
differentiates the coefficients in ``a_j`` (``qbx.py:130-139``).

The tau-derivatives are taken by mpmath's numerical Taylor expansion at 30
digits of the line function, whose radial parts are the closed forms of the
kernels (``kernel.py:639-897``: log r, 1/r, H0(k r), e^{ikr}/r,
H0(i lam r), e^{-lam r}/r) and of their first two r-derivatives -- an
evaluation route independent of the product's power-series recurrences.
"""

from __future__ import annotations

import mpmath as mp
import numpy as np

KINDS = ("lap2", "lap3", "helm2", "helm3", "yuk2", "yuk3")

SCALE = {
    "lap2": -1/(2*np.pi), "lap3": 1/(4*np.pi), "helm2": 0.25j, "helm3": 1/(4*np.pi),
    "yuk2": 0.25j, "yuk3": 1/(4*np.pi),
}


def _radial(kind, k):
    """G, G', G'' of the kernel's radial expression (no scaling constant)."""
    if kind == "lap2":
        return (lambda r: mp.log(r), lambda r: 1/r, lambda r: -1/r**2)
    if kind == "lap3":
        return (lambda r: 1/r, lambda r: -1/r**2, lambda r: 2/r**3)
    if kind in ("helm3", "yuk3"):
        s = 1j*k if kind == "helm3" else -k
        return (lambda r: mp.exp(s*r)/r,
                lambda r: mp.exp(s*r)*(s*r - 1)/r**2,
                lambda r: mp.exp(s*r)*(s**2/r - 2*s/r**2 + 2/r**3))
    s = k if kind == "helm2" else 1j*k
    return (lambda r: mp.hankel1(0, s*r),
            lambda r: -s*mp.hankel1(1, s*r),
            lambda r: -s**2*(mp.hankel1(0, s*r) - mp.hankel1(1, s*r)/(s*r)))


def line_kernel(kind, k, a, b, nrm=None, tgt_axis=-1):
    """tau -> D_s K(a + tau b) (with the a_j derivative when tgt_axis >= 0)."""
    G, G1, G2 = _radial(kind, k)
    a = [mp.mpf(float(x)) for x in a]
    b = [mp.mpf(float(x)) for x in b]
    n = None if nrm is None else [mp.mpf(float(x)) for x in nrm]

    def f(tau):
        d = [ai + tau*bi for ai, bi in zip(a, b)]
        r = mp.sqrt(sum(x*x for x in d))
        if n is None:
            if tgt_axis < 0:
                return G(r)
            return d[tgt_axis]*G1(r)/r
        nd = sum(ni*di for ni, di in zip(n, d))
        if tgt_axis < 0:
            return -nd*G1(r)/r
        j = tgt_axis
        return -n[j]*G1(r)/r - nd*d[j]*(G2(r) - G1(r)/r)/r**2
    return f


def pair_value(kind, k, order, a, b, nrm=None, tgt_axis=-1):
    """sum_{i<=order} (1/i!) d^i/dtau^i D_s K(a + tau b) at 0 (no scaling)."""
    with mp.workdps(30):
        coeffs = mp.taylor(line_kernel(kind, k, a, b, nrm, tgt_axis), 0, order)
        return complex(sum(coeffs))


def layer_potential(kind, k, order, targets, centers, sources, density, normals=None,
                    tgt_axis=-1, src_axis=-1):
    """Dense line-Taylor layer potential ``result[t] = scale sum_s sigma_s
    value(t, s)``; ``normals`` for the directional source derivative, or
    ``src_axis`` >= 0 for the axis source derivative."""
    targets, centers, sources = (np.asarray(x, np.float64) for x in (targets, centers, sources))
    dim, nt = targets.shape
    ns = sources.shape[1]
    out = np.zeros(nt, np.complex128)
    for it in range(nt):
        c = centers[:, it]
        b = targets[:, it] - c
        acc = 0j
        for js in range(ns):
            a = c - sources[:, js]
            nrm = None
            if normals is not None:
                nrm = normals[:, js]
            elif src_axis >= 0:
                nrm = np.eye(dim)[src_axis]
            acc += density[js]*pair_value(kind, k, order, a, b, nrm, tgt_axis)
        out[it] = SCALE[kind]*acc
    return out


def direct_value(kind, k, d, nrm=None, tgt_axis=-1):
    """D_s K(d) itself (the limit of the line-Taylor value as the order grows
    when |b| < |a|)."""
    with mp.workdps(40):
        return complex(line_kernel(kind, k, d, np.zeros_like(d), nrm, tgt_axis)(0))


__all__ = ["KINDS", "SCALE", "direct_value", "layer_potential", "line_kernel", "pair_value"]
