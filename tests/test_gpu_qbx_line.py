"""GPU parity of the line-Taylor layer potential for the Laplace, Helmholtz
and Yukawa kernels in 2D and 3D (``sumpyhip_qbx_line``,
``sumpyhip_qbx_line_matrix``) against the mpmath oracle
(``oracle/qbx_line.py``), the matrix generators against the applier, and the
reference's double-layer jump test (``sumpy/test/test_kernels.py:842-905``:
DLP of a unit density on an ellipse (2D, 1600 nodes, QBX radius 0.02) and a
torus (3D, 200 x 200 nodes, radius 0.25), order 3, interior minus exterior
limit = -1 within 3e-7 resp. 1e-3).
"""

import numpy as np
import pytest
import torch

from oracle import qbx_line as ql

pytestmark = pytest.mark.gpu

KIND_ID = {"lap2": 0, "lap3": 1, "helm2": 2, "helm3": 3, "yuk2": 4, "yuk3": 5}
K = {"lap2": 0.0, "lap3": 0.0, "helm2": 1.3, "helm3": 1.3, "yuk2": 0.9, "yuk3": 0.9}
# source kind, source axis, target axis
MODES = {"slp": (0, -1, -1), "dlp": (1, -1, -1), "dsrc1": (2, 1, -1), "slp_dx": (0, -1, 0),
         "dlp_dy": (1, -1, 1)}


def _case(dim, seed, nt=3, ns=5):
    rng = np.random.default_rng(seed)
    src = rng.standard_normal((dim, ns))
    ctr = 2.5 + 0.3*rng.standard_normal((dim, nt))
    tgt = ctr + 0.25*rng.standard_normal((dim, nt))
    nrm = rng.standard_normal((dim, ns))
    nrm /= np.linalg.norm(nrm, axis=0)
    dens = rng.standard_normal(ns)
    return tgt, ctr, src, nrm, dens


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def _call_line(kind, order, mode, tgt, ctr, src, nrm, dens, accumulate=0, out=None, k=None):
    from sumpy_amd import _capi
    skind, sax, tax = MODES[mode]
    dim = tgt.shape[0]
    t, c, s, n = ([_dev(x[i]) for i in range(dim)] + [None]*(3 - dim)
                  for x in (tgt, ctr, src, nrm))
    d = _dev(dens)
    nt = tgt.shape[1]
    if out is None:
        out = torch.zeros((nt, 2), dtype=torch.float64, device="cuda")
    sc = complex(ql.SCALE[kind])
    print(f"qbx_line {kind} order {order} mode {mode} nt {nt} ns {src.shape[1]}", flush=True)
    k = complex(K[kind] if k is None else k)
    _capi.call("sumpyhip_qbx_line", KIND_ID[kind], order, skind, sax, tax, k.real, k.imag, nt,
               *[_capi.ptr(x) for x in t], *[_capi.ptr(x) for x in c], src.shape[1],
               *[_capi.ptr(x) for x in s], *[_capi.ptr(x) for x in n], _capi.ptr(d),
               sc.real, sc.imag, accumulate, _capi.ptr(out), 0)
    torch.cuda.synchronize()
    return torch.view_as_complex(out).cpu().numpy()


@pytest.mark.parametrize("kind", ql.KINDS)
@pytest.mark.parametrize("mode", list(MODES))
def test_qbx_line_matches_oracle(kind, mode):
    dim = 3 if kind.endswith("3") else 2
    tgt, ctr, src, nrm, dens = _case(dim, 3 + KIND_ID[kind])
    order = 4
    got = _call_line(kind, order, mode, tgt, ctr, src, nrm, dens)
    skind, sax, tax = MODES[mode]
    ref = ql.layer_potential(kind, K[kind], order, tgt, ctr, src, dens,
                             normals=nrm if skind == 1 else None, tgt_axis=tax,
                             src_axis=sax if skind == 2 else -1)
    err = np.abs(got - ref).max()/np.abs(ref).max()
    print(f"{kind} {mode}: max rel. err {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("kind", ["lap3", "helm2", "yuk3"])
def test_qbx_line_high_order_matches_oracle(kind):
    dim = 3 if kind.endswith("3") else 2
    tgt, ctr, src, nrm, dens = _case(dim, 21, nt=2, ns=3)
    got = _call_line(kind, 12, "dlp", tgt, ctr, src, nrm, dens)
    ref = ql.layer_potential(kind, K[kind], 12, tgt, ctr, src, dens, normals=nrm)
    assert np.abs(got - ref).max() <= 1e-10*np.abs(ref).max()


def test_qbx_line_accumulates_and_rejects_bad_input():
    from sumpy_amd import _capi
    tgt, ctr, src, nrm, dens = _case(3, 5)
    a = _call_line("helm3", 3, "slp", tgt, ctr, src, nrm, dens)
    out = torch.view_as_real(torch.as_tensor(a, device="cuda")).contiguous()
    b = _call_line("helm3", 3, "slp", tgt, ctr, src, nrm, dens, accumulate=1, out=out)
    assert np.allclose(b, 2*a, rtol=1e-15, atol=0)
    with pytest.raises(ValueError, match="order 13"):
        _capi.call("sumpyhip_qbx_line", 3, 13, 0, -1, -1, 1.0, 0.0, 1, *[None]*6, 0,
                   *[None]*6, None, 1.0, 0.0, 0, None, 0)


@pytest.mark.parametrize("mode", list(MODES))
def test_qbx_line_helm2_complex_k_matches_oracle(mode):
    """2D Helmholtz with a complex (damped) k: H_n of complex argument."""
    k = 1.3 + 0.4j
    tgt, ctr, src, nrm, dens = _case(2, 31, nt=2, ns=3)
    got = _call_line("helm2", 6, mode, tgt, ctr, src, nrm, dens, k=k)
    skind, sax, tax = MODES[mode]
    ref = ql.layer_potential("helm2", k, 6, tgt, ctr, src, dens,
                             normals=nrm if skind == 1 else None, tgt_axis=tax,
                             src_axis=sax if skind == 2 else -1)
    err = np.abs(got - ref).max()/np.abs(ref).max()
    print(f"helm2 k={k} {mode}: max rel. err {err:.2e}")
    assert err < 1e-12


def test_layer_potential_helm2_complex_k(actx):
    """LayerPotential with HelmholtzKernel(2, allow_evanescent=True) and the
    reference's evanescent k (``sumpy/test/test_kernels.py:168``); without
    allow_evanescent a complex k is refused."""
    from sumpy_amd.expansion.local import LineTaylorLocalExpansion
    from sumpy_amd.kernel import DirectionalSourceDerivative, HelmholtzKernel
    from sumpy_amd.qbx import LayerPotential
    k = 0.2*(0.707 + 0.707j)
    tgt, ctr, src, nrm, dens = _case(2, 12, nt=4, ns=6)
    T, C, S, N = (_dev(x) for x in (tgt, ctr, src, nrm))
    for evanescent in (True, False):
        base = HelmholtzKernel(2, allow_evanescent=evanescent)
        lp = LayerPotential(LineTaylorLocalExpansion(base, 5),
                            source_kernels=[DirectionalSourceDerivative(base)],
                            target_kernels=[base], value_dtypes=np.complex128)
        if not evanescent:
            with pytest.raises(ValueError, match="allow_evanescent"):
                lp(actx, T, S, C, [_dev(dens)], _dev(np.full(4, 0.3)), src_derivative_dir=N,
                   k=k)
            continue
        got, = lp(actx, T, S, C, [_dev(dens)], _dev(np.full(4, 0.3)), src_derivative_dir=N,
                  k=k)
        ref = ql.layer_potential("helm2", k, 5, tgt, ctr, src, dens, normals=nrm)
        assert np.abs(got.cpu().numpy() - ref).max() <= 1e-12*np.abs(ref).max()


def _layer_potential(kind, dim):
    from sumpy_amd.expansion.local import LineTaylorLocalExpansion
    from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel, YukawaKernel
    base = {"lap": LaplaceKernel, "helm": HelmholtzKernel, "yuk": YukawaKernel}[kind[:-1]](dim)
    return base, LineTaylorLocalExpansion


@pytest.mark.parametrize("kind", ["lap3", "helm2", "helm3", "yuk2", "yuk3"])
def test_layer_potential_and_matrices(actx, kind):
    """LayerPotential (SLP + DLP with strength_usage, SLP target derivative)
    for the generic kernels, and the matrix generators equal to it."""
    from sumpy_amd.kernel import AxisTargetDerivative, DirectionalSourceDerivative
    from sumpy_amd.qbx import (
        LayerPotential,
        LayerPotentialMatrixGenerator,
        LayerPotentialMatrixSubsetGenerator,
    )
    dim = 3 if kind.endswith("3") else 2
    base, Expn = _layer_potential(kind, dim)
    tgt, ctr, src, nrm, dens = _case(dim, 9, nt=4, ns=6)
    kw = {"k": K[kind]} if kind.startswith("helm") else (
        {"lam": K[kind]} if kind.startswith("yuk") else {})
    skn = [base, DirectionalSourceDerivative(base)]
    tkn = [base, AxisTargetDerivative(0, base)]
    lp = LayerPotential(Expn(base, 5), source_kernels=skn, target_kernels=tkn,
                        strength_usage=[0, 0], value_dtypes=np.complex128)
    T, C, S, N = (_dev(x) for x in (tgt, ctr, src, nrm))
    got = lp(actx, T, S, C, [_dev(dens)], _dev(np.full(4, 0.3)), src_derivative_dir=N, **kw)
    for ti, ax in enumerate((-1, 0)):
        ref = sum(ql.layer_potential(kind, K[kind], 5, tgt, ctr, src, dens, normals=n,
                                     tgt_axis=ax) for n in (None, nrm))
        g = got[ti].cpu().numpy()
        assert np.abs(g - ref).max() <= 1e-12*np.abs(ref).max(), (kind, ax)
    mat, = LayerPotentialMatrixGenerator(Expn(base, 5), source_kernels=skn,
                                        target_kernels=[base])(
        actx, T, S, C, _dev(np.full(4, 0.3)), src_derivative_dir=N, **kw)
    m = mat.cpu().numpy()
    assert m.shape == (4, 6)
    assert np.allclose(m @ dens, got[0].cpu().numpy(), rtol=1e-12, atol=1e-14)
    ti = np.array([0, 3, 2, 1, 3], np.int32)
    si = np.array([5, 0, 2, 2, 4], np.int32)
    sub, = LayerPotentialMatrixSubsetGenerator(Expn(base, 5), source_kernels=skn,
                                              target_kernels=[base])(
        actx, T, S, C, _dev(np.full(4, 0.3)), torch.as_tensor(ti, device="cuda"),
        torch.as_tensor(si, device="cuda"), src_derivative_dir=N, **kw)
    assert np.array_equal(sub.cpu().numpy(), m[ti, si])


def _ellipse(npoints=1600, a=1.0, b=0.5):
    t = np.linspace(0, 1, npoints, endpoint=False)
    th = 2*np.pi*t
    nodes = np.array([a*np.cos(th), b*np.sin(th)])
    dx, dy = -2*np.pi*a*np.sin(th), 2*np.pi*b*np.cos(th)
    speed = np.hypot(dx, dy)
    return nodes, np.array([dy, -dx])/speed, np.full(npoints, 1.0/npoints), speed


def _torus(rmaj=1.0, rmin=0.5, nmaj=200, nmin=200):
    u = np.linspace(0, 2*np.pi, nmaj, endpoint=False)
    v = np.linspace(0, 2*np.pi, nmin, endpoint=False)
    u, v = np.meshgrid(u, v)
    A = rmaj + rmin*np.cos(v)
    nodes = np.array([np.cos(u)*A, np.sin(u)*A, rmin*np.sin(v)])
    nrm = np.array([np.cos(u)*np.cos(v), np.sin(u)*np.cos(v), np.sin(v)])
    area = A*rmin
    w = np.full(area.shape, (2*np.pi)**2/(nmaj*nmin))
    return nodes.reshape(3, -1), nrm.reshape(3, -1), w.reshape(-1), area.reshape(-1)


@pytest.mark.parametrize("kind", ["lap", "helm", "yuk"])
@pytest.mark.parametrize("dim", [2, 3])
def test_dlp_jump(actx, kind, dim):
    from sumpy_amd.kernel import DirectionalSourceDerivative
    from sumpy_amd.qbx import LayerPotential
    base, Expn = _layer_potential(kind + str(dim), dim)
    if dim == 2:
        nodes, nrm, w, area = _ellipse()
        radius, tol = 0.02, 3e-7
    else:
        nodes, nrm, w, area = _torus()
        radius, tol = 0.25, 1e-3
    # both targets at node 0, centres inside (-1) and outside (+1)
    tgt = nodes[:, :1] + np.zeros(2)
    ctr = nodes[:, :1] + np.array([-1.0, 1.0])*radius*nrm[:, :1]
    kw = {"k": 1.0} if kind == "helm" else ({"lam": 1.0} if kind == "yuk" else {})
    lp = LayerPotential(Expn(base, 3), source_kernels=(DirectionalSourceDerivative(base),),
                        target_kernels=(base,), value_dtypes=np.complex128)
    print(f"jump {kind}{dim}: {nodes.shape[1]} nodes", flush=True)
    y, = lp(actx, _dev(tgt), _dev(nodes), _dev(ctr), [_dev(w*area)],
            _dev(np.full(2, radius)), src_derivative_dir=_dev(nrm), **kw)
    inside, outside = y.cpu().numpy()
    err = abs((inside - outside) + 1)
    print(f"jump {kind}{dim}: inside - outside = {inside - outside}, err {err:.2e}")
    assert err < tol
