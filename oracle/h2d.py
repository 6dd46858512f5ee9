"""CPU oracle for the 2D Hankel/Bessel expansions (H2D, Y2D) -- TEST
INFRASTRUCTURE ONLY (imported by ``tests/`` alone).

Every function takes the Bessel argument scale ``k``: the Helmholtz
wavenumber (H2D), or ``1j*lam`` for the Yukawa kernel (Y2D), whose
expansions are the H2D formulas at the scaled argument
(``sumpy/expansion/multipole.py:551-567``, ``local.py:683-704``) and whose
Green's function is (i/4) H0(i lam r) = K0(lam r)/(2 pi)
(``sumpy/kernel.py:866-897``); scipy evaluates ``jv`` and ``hankel1`` at
complex arguments.

A numpy/scipy restatement of the reference's formulas, one function per
operation, each citing the reference file:line it follows.  Coefficients
are indexed ``c + order`` for ``c = -order..order``
(``sumpy/expansion/multipole.py:431-438``).  Vectors follow the reference:
``avec = center - source`` (``sumpy/qbx.py:284``), ``bvec = target -
center`` (``qbx.py:285``), ``dvec = tgt_center - src_center``
(``sumpy/e2e.py:229``).  The Bessel and Hankel functions are scipy's
(``scipy.special.jv``, ``hankel1``), not this build's device routines.

G(r) = (i/4) H0^(1)(k r) (``sumpy/kernel.py:754-844`` with the 2D
expression and scaling); the expansions carry no scaling, the evaluations
apply it (``get_global_scaling_const``).
"""

from __future__ import annotations

import numpy as np
from scipy import special

SCALE = 0.25j


def _ids(order):
    return np.arange(-order, order + 1)


def p2p(k, targets, sources, weights, exclude_self_t2s=None, dipoles=None, grad=False):
    """Direct sum (i/4) H0(k |t - s|) q_s; pairs at zero distance add nothing.

    dipoles = (q n_x, q n_y): ``DirectionalSourceDerivative`` sources,
    n . grad_s G(t - s) (``sumpy/kernel.py:2288-2385``), added to the
    charges.  grad: return (potential, d/dt_x, d/dt_y)
    (``AxisTargetDerivative``), from the Hessian G'' d d^T/r^2 +
    G'/r (I - d d^T/r^2) with G' = -(i/4) k H1(k r), G'' = -(i/4) k^2 H1'(k r)
    (scipy's ``h1vp``)."""
    t = np.asarray(targets, np.float64)
    s = np.asarray(sources, np.float64)
    dx = t[0][:, None] - s[0][None, :]
    dy = t[1][:, None] - s[1][None, :]
    r = np.hypot(dx, dy)
    nz = r > 0
    if exclude_self_t2s is not None:
        nz[np.arange(len(exclude_self_t2s)), exclude_self_t2s] = False
    rr = np.where(nz, r, 1.0)
    g = np.where(nz, SCALE*special.hankel1(0, k*rr), 0)
    g1 = np.where(nz, -SCALE*k*special.hankel1(1, k*rr), 0)        # G'
    g2 = np.where(nz, -SCALE*k*k*special.h1vp(1, k*rr), 0)         # G''
    q = np.asarray(weights, np.complex128)
    ux, uy = dx/rr, dy/rr
    gx, gy = g1*ux, g1*uy                                          # grad_t G
    pot = g @ q
    hxx = g2*ux*ux + g1/rr*(1 - ux*ux)
    hyy = g2*uy*uy + g1/rr*(1 - uy*uy)
    hxy = (g2 - g1/rr)*ux*uy
    outx, outy = gx @ q, gy @ q
    if dipoles is not None:
        wx, wy = (np.asarray(w, np.complex128) for w in dipoles)
        # n . grad_s G(t - s) = -n . grad_t G; its target gradient -H n
        pot = pot - (gx @ wx + gy @ wy)
        outx = outx - (hxx @ wx + hxy @ wy)
        outy = outy - (hxy @ wx + hyy @ wy)
    return (pot, outx, outy) if grad else pot


def _basis(order, k, vec, sign, hankel, rscale_pow):
    """Values and x/y derivatives (with respect to the components of vec) of
    F_c(v) = Z_c(k|v|) e^{i sign c theta_v} rscale_pow(c), c = -order..order,
    Z = H^(1) or J, by the chain rule in polar coordinates (d/drho through
    scipy's ``h1vp`` / ``jvp``, d/dtheta = i sign c)."""
    c = _ids(order)[:, None]
    v = np.asarray(vec, np.float64)
    rho, th = np.hypot(v[0], v[1])[None, :], np.arctan2(v[1], v[0])[None, :]
    z, zp = ((special.hankel1(c, k*rho), special.h1vp(c, k*rho)) if hankel
             else (special.jv(c, k*rho), special.jvp(c, k*rho)))
    ph = np.exp(1j*sign*c*th)*rscale_pow(c)
    f = z*ph
    frho, fth = k*zp*ph, 1j*sign*c*f
    fx = np.cos(th)*frho - np.sin(th)/rho*fth
    fy = np.sin(th)*frho + np.cos(th)/rho*fth
    return f, fx, fy


def p2m(order, k, center, sources, weights, rscale, dipoles=None):
    """``HankelBased2DMultipoleExpansion.coefficients_from_source``
    (``sumpy/expansion/multipole.py:441-470``): J_c(k|a|) / rscale^|c|
    e^{-i c theta_s}, theta_s = atan2(-a_y, -a_x), summed with the weights;
    dipoles = (q n_x, q n_y) add n . grad_s of the same terms."""
    if dipoles is not None:
        v = np.asarray(sources, np.float64) - np.asarray(center, np.float64)[:, None]
        f, fx, fy = _basis(order, k, v, -1, False, lambda c: rscale**(-np.abs(c)))
        return (f @ np.asarray(weights, np.complex128) + fx @ np.asarray(dipoles[0], complex)
                + fy @ np.asarray(dipoles[1], complex))
    c = _ids(order)
    a = np.asarray(center, np.float64)[:, None] - np.asarray(sources, np.float64)
    rho = np.hypot(a[0], a[1])
    th = np.arctan2(-a[1], -a[0])
    terms = (special.jv(c[:, None], k*rho[None, :]) / rscale**np.abs(c)[:, None]
             * np.exp(1j*c[:, None]*(-th[None, :])))
    return terms @ np.asarray(weights, np.complex128)


def m2p(order, k, center, coeffs, targets, rscale, grad=False):
    """``HankelBased2DMultipoleExpansion.evaluate`` (``multipole.py:472-497``)
    times the kernel scaling: sum_c M_c H_c(k|b|) rscale^|c| e^{i c theta_b};
    grad: (potential, d/dx, d/dy) at the targets."""
    if grad:
        v = np.asarray(targets, np.float64) - np.asarray(center, np.float64)[:, None]
        f, fx, fy = _basis(order, k, v, 1, True, lambda c: rscale**np.abs(c))
        cf = np.asarray(coeffs)
        return SCALE*(cf @ f), SCALE*(cf @ fx), SCALE*(cf @ fy)
    c = _ids(order)
    b = np.asarray(targets, np.float64) - np.asarray(center, np.float64)[:, None]
    rho = np.hypot(b[0], b[1])
    th = np.arctan2(b[1], b[0])
    terms = (special.hankel1(c[:, None], k*rho[None, :]) * rscale**np.abs(c)[:, None]
             * np.exp(1j*c[:, None]*th[None, :]))
    return SCALE*(np.asarray(coeffs) @ terms)


def m2m(src_order, tgt_order, k, coeffs, src_center, tgt_center, src_rscale, tgt_rscale):
    """``HankelBased2DMultipoleExpansion.translate_from``
    (``multipole.py:499-530``): new_j = sum_m M_m J_{m-j}(k|d|)
    src_rscale^|m| / tgt_rscale^|j| e^{i (m-j) theta_d}."""
    d = np.asarray(tgt_center, np.float64) - np.asarray(src_center, np.float64)
    dl, th = np.hypot(d[0], d[1]), np.arctan2(d[1], d[0])
    m, j = _ids(src_order), _ids(tgt_order)
    nm = m[None, :] - j[:, None]
    T = (special.jv(nm, k*dl) * src_rscale**np.abs(m)[None, :]
         / tgt_rscale**np.abs(j)[:, None] * np.exp(1j*nm*th))
    return T @ np.asarray(coeffs)


def p2l(order, k, center, sources, weights, rscale, dipoles=None):
    """``FourierBesselLocalExpansionMixin.coefficients_from_source``
    (``sumpy/expansion/local.py:560-583``): H_c(k|a|) rscale^|c|
    e^{i c theta_s}, theta_s = atan2(-a_y, -a_x); dipoles as in p2m."""
    if dipoles is not None:
        v = np.asarray(sources, np.float64) - np.asarray(center, np.float64)[:, None]
        f, fx, fy = _basis(order, k, v, 1, True, lambda c: rscale**np.abs(c))
        return (f @ np.asarray(weights, np.complex128) + fx @ np.asarray(dipoles[0], complex)
                + fy @ np.asarray(dipoles[1], complex))
    c = _ids(order)
    a = np.asarray(center, np.float64)[:, None] - np.asarray(sources, np.float64)
    rho = np.hypot(a[0], a[1])
    th = np.arctan2(-a[1], -a[0])
    terms = (special.hankel1(c[:, None], k*rho[None, :]) * rscale**np.abs(c)[:, None]
             * np.exp(1j*c[:, None]*th[None, :]))
    return terms @ np.asarray(weights, np.complex128)


def l2p(order, k, center, coeffs, targets, rscale, grad=False):
    """``FourierBesselLocalExpansionMixin.evaluate`` (``local.py:585-605``)
    times the kernel scaling: sum_c L_c J_c(k|b|) / rscale^|c|
    e^{-i c theta_b}; grad as in m2p."""
    if grad:
        v = np.asarray(targets, np.float64) - np.asarray(center, np.float64)[:, None]
        f, fx, fy = _basis(order, k, v, -1, False, lambda c: rscale**(-np.abs(c)))
        cf = np.asarray(coeffs)
        return SCALE*(cf @ f), SCALE*(cf @ fx), SCALE*(cf @ fy)
    c = _ids(order)
    b = np.asarray(targets, np.float64) - np.asarray(center, np.float64)[:, None]
    rho = np.hypot(b[0], b[1])
    th = np.arctan2(b[1], b[0])
    terms = (special.jv(c[:, None], k*rho[None, :]) / rscale**np.abs(c)[:, None]
             * np.exp(-1j*c[:, None]*th[None, :]))
    return SCALE*(np.asarray(coeffs) @ terms)


def l2l(src_order, tgt_order, k, coeffs, src_center, tgt_center, src_rscale, tgt_rscale):
    """``FourierBesselLocalExpansionMixin.translate_from`` from a local
    (``local.py:607-640``): new_j = sum_m L_m J_{m-j}(k|d|) tgt_rscale^|j|
    / src_rscale^|m| e^{-i (m-j) theta_d}."""
    d = np.asarray(tgt_center, np.float64) - np.asarray(src_center, np.float64)
    dl, th = np.hypot(d[0], d[1]), np.arctan2(d[1], d[0])
    m, j = _ids(src_order), _ids(tgt_order)
    nm = m[None, :] - j[:, None]
    T = (special.jv(nm, k*dl) / src_rscale**np.abs(m)[None, :]
         * tgt_rscale**np.abs(j)[:, None] * np.exp(-1j*nm*th))
    return T @ np.asarray(coeffs)


def m2l_class_data(src_order, tgt_order, k, dvec):
    """``FourierBesselM2LTranslation.translation_classes_dependent_data``
    (``sumpy/expansion/m2l.py:1045-1080``): entry ``n + src_order +
    tgt_order`` is H_n(k|d|) e^{i n theta_d}, n = -(src+tgt)..(src+tgt)."""
    d = np.asarray(dvec, np.float64)
    dl, th = np.hypot(d[0], d[1]), np.arctan2(d[1], d[0])
    n = np.arange(-(src_order + tgt_order), src_order + tgt_order + 1)
    return special.hankel1(n, k*dl)*np.exp(1j*n*th)


def m2l(src_order, tgt_order, k, coeffs, dvec, src_rscale, tgt_rscale, data=None):
    """``FourierBesselM2LTranslation.translate`` with its pre- and
    postprocessing (``m2l.py:1008-1115``): M'_m = M_m src_rscale^|m|,
    L_j = sum_m data[m + j + src + tgt] M'_m, then L_j tgt_rscale^|j| (-1)^j."""
    if data is None:
        data = m2l_class_data(src_order, tgt_order, k, dvec)
    m, j = _ids(src_order), _ids(tgt_order)
    pre = np.asarray(coeffs)*src_rscale**np.abs(m)
    T = data[m[None, :] + j[:, None] + src_order + tgt_order]
    return (T @ pre)*tgt_rscale**np.abs(j)*(-1.0)**np.abs(j)


__all__ = ["SCALE", "l2l", "l2p", "m2l", "m2l_class_data", "m2m", "m2p", "p2l", "p2m",
           "p2p"]
