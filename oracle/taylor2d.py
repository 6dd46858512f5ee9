"""CPU oracle for the 2D volume-Taylor expansions (full and PDE-compressed)
of the Laplace and Helmholtz kernels -- TEST INFRASTRUCTURE ONLY (imported by
``tests/`` alone).

A numpy restatement of the reference's formulas, one function per
operation:

* P2M: full moments ``sum_s q_s (a/rho)^m / m!``, ``a = centre - source``
  (``sumpy/expansion/multipole.py:78-114``, ``expansion/loopy.py:183-188``),
  compressed by the transposed projection (``expansion/__init__.py:588-597``);
* M2P: ``scale sum_m M_m rho^|m| d^m G(t - c)`` (``multipole.py:131-157``);
* P2L: ``L_m = sum_s q_s rho^|m| d^m G(c - s)`` (``local.py:226-271``);
* L2P: ``scale sum_m (P L)_m (b/rho)^m / m!`` (``local.py:284-309``);
* M2M, L2L: Taylor re-expansion of the moments / of the local polynomial
  (the reference's "slow" binomial forms, ``multipole.py:363-402``,
  ``local.py:468-490``);
* M2L: ``L_t = sum_s M_s rho^|t+s| d^(t+s) G(d)`` (``m2l.py:363-389``).

The projection P is ``oracle.tables.Wrangler(order, 2, ...).projection``; the
kernel derivatives come from the radial form G(r) = h(r^2/2) (Hermite-type
expansion of the Cartesian derivatives, h^(n) from DLMF 10.6.6 and scipy's
Hankel functions) -- independent of the product's W-operator recurrences.
Scaling constants: ``-1/(2 pi)`` (Laplace),
``i/4`` (Helmholtz), applied by the evaluations only.  The Helmholtz k may be
complex (scipy's Hankel functions and the projection take complex k).
"""

from __future__ import annotations

from math import factorial

import numpy as np

from . import tables

SCALE = {"lap": -1/(2*np.pi), "helm": 0.25j}


def _hder(kind, k, r, n):
    """h^(n)(s) for G(r) = h(s), s = r^2/2: (1/r d/dr)^n G (DLMF 10.6.6 with
    nu = 0: (1/z d/dz)^n H0(z) = (-1)^n z^-n H_n(z); log r = log(2 s)/2)."""
    if kind == "lap":
        if n == 0:
            return np.log(r)
        return 0.5*(-1)**(n - 1)*factorial(n - 1)*(r*r/2)**(-n)
    from scipy.special import hankel1
    return (-k)**n*r**(-n)*hankel1(n, k*r)


def deriv(kind, k, d, mx, my):
    """d_x^mx d_y^my G at the vector d (no scaling constant), from
    d_x^a f(x^2/2) = sum_i a!/(2^i i! (a-2i)!) x^(a-2i) f^(a-i):
    d_x^mx d_y^my h(s) = sum_{i, j} c_i(mx) c_j(my) x^(mx-2i) y^(my-2j)
    h^(mx+my-i-j)(s)."""
    x, y = float(d[0]), float(d[1])
    r = np.hypot(x, y)

    def c(a, i):
        return factorial(a)/(2**i*factorial(i)*factorial(a - 2*i))
    v = 0j
    for i in range(mx//2 + 1):
        for j in range(my//2 + 1):
            v += (c(mx, i)*c(my, j)*x**(mx - 2*i)*y**(my - 2*j)
                  * _hder(kind, k, r, mx + my - i - j))
    return complex(v)


def _wr(order, compressed, kind):
    return tables.wrangler(order, 2, compressed, helmholtz=kind == "helm")


def _proj(order, compressed, kind, rscale, k):
    if kind != "helm":
        return _wr(order, compressed, kind).projection(rscale, 0.0)
    return _wr(order, compressed, kind).projection(
        rscale, complex(k) if np.iscomplexobj(k) and complex(k).imag else float(np.real(k)))


def _mono(mi, v, rscale):
    return np.prod([(v[i]/rscale)**mi[i]/factorial(mi[i]) for i in range(2)])


def _mono_d(mi, v, rscale, ax):
    """d/dv_ax of (v/rscale)^mi/mi!."""
    if mi[ax] == 0:
        return 0.0
    m = list(mi)
    m[ax] -= 1
    return _mono(m, v, rscale)/rscale


def p2m(order, compressed, kind, k, center, sources, weights, rscale, dipoles=None):
    """dipoles = (q n_x, q n_y): n . grad_s of the moments (a = c - s, so
    -n . grad_a)."""
    w = _wr(order, compressed, kind)
    srcs = np.asarray(sources).T
    full = np.array([sum(q*_mono(mi, np.asarray(center) - s, rscale)
                         for s, q in zip(srcs, weights))
                     for mi in w.full], dtype=np.complex128)
    if dipoles is not None:
        full -= np.array([sum(dx*_mono_d(mi, np.asarray(center) - s, rscale, 0)
                              + dy*_mono_d(mi, np.asarray(center) - s, rscale, 1)
                              for s, dx, dy in zip(srcs, *dipoles))
                          for mi in w.full], dtype=np.complex128)
    return _proj(order, compressed, kind, rscale, k).T @ full


def m2p(order, compressed, kind, k, center, coeffs, targets, rscale, grad=False):
    """grad: (potential, d/dx, d/dy): rho^|m| d^(m + e) G."""
    w = _wr(order, compressed, kind)
    out = []
    for t in np.asarray(targets).T:
        b = t - np.asarray(center)
        out.append([sum(c*rscale**sum(mi)*deriv(kind, k, b, mi[0] + ex, mi[1] + ey)
                        for mi, c in zip(w.stored, coeffs))
                    for ex, ey in ((0, 0), (1, 0), (0, 1))[:3 if grad else 1]])
    out = SCALE[kind]*np.array(out)
    return tuple(out.T) if grad else out[:, 0]


def p2l(order, compressed, kind, k, center, sources, weights, rscale, dipoles=None):
    """dipoles as in p2m: n . grad_s of rho^|m| d^m G(c - s)."""
    w = _wr(order, compressed, kind)
    srcs = np.asarray(sources).T
    out = np.array([sum(q*rscale**sum(mi)*deriv(kind, k, np.asarray(center) - s, *mi)
                        for s, q in zip(srcs, weights))
                    for mi in w.stored], dtype=np.complex128)
    if dipoles is not None:
        out -= np.array([sum(rscale**sum(mi)*(
            dx*deriv(kind, k, np.asarray(center) - s, mi[0] + 1, mi[1])
            + dy*deriv(kind, k, np.asarray(center) - s, mi[0], mi[1] + 1))
            for s, dx, dy in zip(srcs, *dipoles)) for mi in w.stored], dtype=np.complex128)
    return out


def l2p(order, compressed, kind, k, center, coeffs, targets, rscale, grad=False):
    """grad: (potential, d/dx, d/dy) of the local polynomial."""
    w = _wr(order, compressed, kind)
    full = _proj(order, compressed, kind, rscale, k) @ np.asarray(coeffs)
    ts = np.asarray(targets).T
    out = np.array([sum(c*_mono(mi, t - np.asarray(center), rscale) for mi, c in zip(w.full, full))
                    for t in ts])
    if not grad:
        return SCALE[kind]*out
    g = [np.array([sum(c*_mono_d(mi, t - np.asarray(center), rscale, ax)
                       for mi, c in zip(w.full, full)) for t in ts]) for ax in (0, 1)]
    return SCALE[kind]*out, SCALE[kind]*g[0], SCALE[kind]*g[1]


def m2m(src_order, tgt_order, compressed, kind, k, coeffs, src_center, tgt_center,
        src_rscale, tgt_rscale):
    """Moments about the new centre from the stored moments (zero elsewhere):
    M'_m = sum_{n<=m} M_n (rs/rt)^|n| ((c_src - c_tgt)/rt)^(m-n)/(m-n)!."""
    ws, wt = _wr(src_order, compressed, kind), _wr(tgt_order, compressed, kind)
    d = np.asarray(tgt_center) - np.asarray(src_center)
    full = []
    for m in wt.full:
        v = 0j
        for n, c in zip(ws.stored, coeffs):
            if n[0] <= m[0] and n[1] <= m[1]:
                v += c*(src_rscale/tgt_rscale)**sum(n)*_mono(
                    (m[0] - n[0], m[1] - n[1]), d, tgt_rscale)
        full.append(v)
    return _proj(tgt_order, compressed, kind, tgt_rscale, k).T @ np.array(full)


def l2l(src_order, tgt_order, compressed, kind, k, coeffs, src_center, tgt_center,
        src_rscale, tgt_rscale):
    """Derivatives of the local polynomial at the new centre:
    L'_m = (rt/rs)^|m| sum_{n>=m} (P L)_n ((c_tgt - c_src)/rs)^(n-m)/(n-m)!."""
    ws, wt = _wr(src_order, compressed, kind), _wr(tgt_order, compressed, kind)
    full = _proj(src_order, compressed, kind, src_rscale, k) @ np.asarray(coeffs)
    d = np.asarray(tgt_center) - np.asarray(src_center)
    out = []
    for m in wt.stored:
        v = 0j
        for n, c in zip(ws.full, full):
            if n[0] >= m[0] and n[1] >= m[1]:
                v += c*_mono((n[0] - m[0], n[1] - m[1]), d, src_rscale)
        out.append(v*(tgt_rscale/src_rscale)**sum(m))
    return np.array(out)


def m2l(src_order, tgt_order, compressed, kind, k, coeffs, dvec, rscale):
    """L_t = sum_s M_s rscale^|t+s| d^(t+s) G(d), d = target centre - source
    centre (same rscale on both sides)."""
    ws, wt = _wr(src_order, compressed, kind), _wr(tgt_order, compressed, kind)
    return np.array([sum(c*rscale**(sum(t) + sum(s))*deriv(kind, k, dvec, t[0] + s[0],
                                                           t[1] + s[1])
                         for s, c in zip(ws.stored, coeffs))
                     for t in wt.stored], dtype=np.complex128)


def stored_ids(order, compressed, kind="lap"):
    return list(_wr(order, compressed, kind).stored)


def p2p(kind, k, targets, sources, weights, dipoles=None, grad=False):
    """Direct sum; dipoles = (q n_x, q n_y) add n . grad_s G(t - s); grad:
    (potential, d/dt_x, d/dt_y).  Helmholtz: ``oracle.h2d.p2p``; Laplace:
    log r with grad log r = d/r^2 and the Hessian I/r^2 - 2 d d^T/r^4."""
    if kind != "lap":
        from . import h2d
        return h2d.p2p(k, targets, sources, weights, dipoles=dipoles, grad=grad)
    t, s = np.asarray(targets), np.asarray(sources)
    dx = t[0][:, None] - s[0][None, :]
    dy = t[1][:, None] - s[1][None, :]
    r2 = dx*dx + dy*dy
    nz = r2 > 0
    rr2 = np.where(nz, r2, 1.0)
    g = np.where(nz, 0.5*np.log(rr2), 0.0)
    gx, gy = np.where(nz, dx/rr2, 0.0), np.where(nz, dy/rr2, 0.0)
    hxx = np.where(nz, 1/rr2 - 2*dx*dx/rr2**2, 0.0)
    hyy = np.where(nz, 1/rr2 - 2*dy*dy/rr2**2, 0.0)
    hxy = np.where(nz, -2*dx*dy/rr2**2, 0.0)
    q = np.asarray(weights)
    pot, ox, oy = g @ q, gx @ q, gy @ q
    if dipoles is not None:
        wx, wy = (np.asarray(w) for w in dipoles)
        pot = pot - (gx @ wx + gy @ wy)
        ox = ox - (hxx @ wx + hxy @ wy)
        oy = oy - (hxy @ wx + hyy @ wy)
    out = SCALE[kind]*pot, SCALE[kind]*ox, SCALE[kind]*oy
    return out if grad else out[0]


__all__ = ["SCALE", "deriv", "l2l", "l2p", "m2l", "m2m", "m2p", "p2l", "p2m", "p2p",
           "stored_ids"]
