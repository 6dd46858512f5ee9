"""Helmholtz-3D pieces of the oracle -- TEST INFRASTRUCTURE ONLY.

Scaled derivatives ``rscale^|m| d^m G`` of ``G = e^{ik|d|}/|d|`` by the
reference's recursion, restated over numpy arrays of points:

* ``RadialDerivativeTaker.diff`` (``sumpy/derivative_taker.py:290-346``):
  ``D(mi, q) = x_i rho D(mi - e_i, q + 1)`` for the first axis with
  ``mi_i == 1``, else ``(mi_i - 1) rho^2 D(mi - 2 e_i, q + 1) + x_i rho
  D(mi - e_i, q + 1)`` for the first axis with ``mi_i >= 2``;
* ``HelmholtzDerivativeTaker.diff`` (``:354-382``) for ``mi = 0``:
  ``g_q = (-(2q - 1) g_{q-1} - k^2 g_{q-2})/r^2`` from
  ``g_0 = e^{ikr}/r`` and ``g_1 = (1/r) d/dr g_0``.

These evaluate the same sums as the GPU's table of chain-rule coefficients
along a different route (memoised recursion over (mi, q)); both are pinned to
sympy's exact differentiation at small orders by the tests.  Pure Python
loops over multi-indices, numpy over points: small cases only.
"""

from __future__ import annotations

import numpy as np

from . import tables


def derivs(x, rscale, k, maxorder):
    """``{mi: rscale^|mi| d^mi G(x)}`` for ``|mi| <= maxorder``; *x* is
    (3, n) (or (3,)), values are complex arrays of shape (n,)."""
    x = np.asarray(x, np.float64)
    if x.ndim == 1:
        x = x[:, None]
    r = np.sqrt((x*x).sum(axis=0))
    g = [np.exp(1j*k*r)/r]
    g.append(np.exp(1j*k*r)*(1j*k*r - 1)/r**3)
    for q in range(2, maxorder + 2):
        g.append((-(2*q - 1)*g[q - 1] - k*k*g[q - 2])/r**2)
    memo = {}

    def d(mi, q):
        key = (mi, q)
        if key in memo:
            return memo[key]
        if mi == (0, 0, 0):
            v = g[q]
        else:
            v = None
            for i in range(3):
                if mi[i] == 1:
                    m1 = list(mi)
                    m1[i] = 0
                    v = x[i]*rscale*d(tuple(m1), q + 1)
                    break
            if v is None:
                for i in range(3):
                    if mi[i] >= 2:
                        m1, m2 = list(mi), list(mi)
                        m1[i] -= 1
                        m2[i] -= 2
                        v = ((mi[i] - 1)*rscale**2*d(tuple(m2), q + 1)
                             + x[i]*rscale*d(tuple(m1), q + 1))
                        break
        memo[key] = v
        return v

    out = {}
    for n in range(maxorder + 1):
        for a in range(n + 1):
            for b in range(n + 1 - a):
                mi = (a, b, n - a - b)
                out[mi] = d(mi, 0)
    return out


def class_data(order_src, order_tgt, vectors, rscale, k, shape):
    """(nclasses, 2, prod(shape)): real and imaginary planes of the scaled
    derivatives at each class vector, row-major in *shape*, zero past
    ``|m| > order_src + order_tgt`` (``sumpy/e2e.py:565-669`` with the
    circulant layout of ``sumpy/expansion/m2l.py:403-537``)."""
    N = order_src + order_tgt
    nc = vectors.shape[1]
    D = derivs(vectors, rscale, k, N)
    out = np.zeros((nc, 2, int(np.prod(shape))))
    for mi, v in D.items():
        if all(mi[i] < shape[i] for i in range(3)):
            off = (mi[0]*shape[1] + mi[1])*shape[2] + mi[2]
            out[:, 0, off] = v.real
            out[:, 1, off] = v.imag
    return out


def m2p(order, target_boxes, src_box_starts, src_box_lists, tree, centers, targets,
        rscale, k, mpoles, src_base_ibox, pot, grad=None, scale=1/(4*np.pi)):
    """``pot[t] += scale sum_s M_s rho^|s| d^s G(t - c)`` (and the target
    gradient) over list 3 (``sumpy/e2p.py:233-343``); complex *mpoles*."""
    ids = [tuple(int(v) for v in mi) for mi in tables.wrangler(order, 3, True).stored]
    bts, btc = tree["box_target_starts"], tree["box_target_counts_nonchild"]
    for i, tb in enumerate(target_boxes):
        sl = slice(int(bts[tb]), int(bts[tb]) + int(btc[tb]))
        for q in range(int(src_box_starts[i]), int(src_box_starts[i + 1])):
            sb = int(src_box_lists[q])
            x = targets[:, sl] - centers[:, sb][:, None]
            D = derivs(x, rscale, k, order + (1 if grad is not None else 0))
            M = mpoles[sb - src_base_ibox]
            pot[sl] += scale*sum(M[j]*D[mi] for j, mi in enumerate(ids))
            if grad is not None:
                for ax in range(3):
                    e = tuple(int(ax == a) for a in range(3))
                    grad[ax][sl] += scale/rscale*sum(
                        M[j]*D[tuple(m + u for m, u in zip(mi, e))]
                        for j, mi in enumerate(ids))
    return pot


def p2l(order, target_boxes, src_box_starts, src_box_lists, tree, centers, sources,
        w0, wd, rscale, k, locs, tgt_base_ibox):
    """``L_t = sum_y [w0 rho^|t| d^t G(c - y) + sum_i w_i rho^|t| d^{t + e_i}
    G(c - y)/rho]`` over list 4, overwriting the box rows
    (``sumpy/p2e.py:268-388``); complex weights."""
    ids = [tuple(int(v) for v in mi) for mi in tables.wrangler(order, 3, True).stored]
    bss, bsc = tree["box_source_starts"], tree["box_source_counts_nonchild"]
    for i, tb in enumerate(target_boxes):
        acc = np.zeros(len(ids), np.complex128)
        for q in range(int(src_box_starts[i]), int(src_box_starts[i + 1])):
            sb = int(src_box_lists[q])
            sl = slice(int(bss[sb]), int(bss[sb]) + int(bsc[sb]))
            x = centers[:, tb][:, None] - sources[:, sl]
            D = derivs(x, rscale, k, order + (1 if wd is not None else 0))
            for j, mi in enumerate(ids):
                v = (w0[sl]*D[mi]).sum()
                if wd is not None:
                    for ax in range(3):
                        e = tuple(int(ax == a) for a in range(3))
                        v += (wd[ax][sl]*D[tuple(m + u for m, u in zip(mi, e))]).sum()/rscale
                acc[j] += v
        locs[tb - tgt_base_ibox] = acc
    return locs
