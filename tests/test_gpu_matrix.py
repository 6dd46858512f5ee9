"""Matrix generators (reference ``sumpy/p2p.py:292-477``, ``sumpy/qbx.py:
342-539``): the full matrix times a strength vector equals the direct sum of
the applier (P2P / LayerPotential); columns equal the oracle's direct sum
(line-Taylor sum for QBX) of a unit source, for every target kernel; the subset generator
returns the full matrix's entries at the requested pairs; excluded self pairs
are zero.  Cases follow the reference's tests (``test_matrixgen.py``):
Laplace 2D/3D and Helmholtz 3D, source and target derivatives, a coordinate
multiplier, and the 2D Laplace QBX SLP/DLP."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b))/np.linalg.norm(np.asarray(b))


def _kernels(name):
    from sumpy_amd.kernel import (
        AxisSourceDerivative,
        AxisTargetDerivative,
        HelmholtzKernel,
        LaplaceKernel,
        TargetPointMultiplier,
    )
    if name == "lap3d":
        k = LaplaceKernel(3)
        return 3, k, [k], [k, AxisTargetDerivative(0, k), TargetPointMultiplier(2, k)], {}
    if name == "lap3d_dipole":
        k = LaplaceKernel(3)
        return 3, k, [AxisSourceDerivative(1, k)], [k, AxisTargetDerivative(2, k)], {}
    if name == "lap2d":
        k = LaplaceKernel(2)
        return 2, k, [k], [k, AxisTargetDerivative(1, k)], {}
    k = HelmholtzKernel(3, allow_evanescent=True)
    return 3, k, [k, AxisSourceDerivative(0, k)], [k, AxisTargetDerivative(1, k)], \
        {"k": 5.0 + 0.5j}


def _oracle_columns(name, tgt, src, j, exclude_self):
    """Column j of each target kernel's matrix of :func:`_kernels`'s case, by
    the oracle's direct sum with unit strength at source j for every source
    kernel (dipole weight -1 along the derivative axis: d/ds_a G(t - s) =
    -d/dt_a G)."""
    n = src.shape[1]
    e = np.zeros(n)
    e[j] = 1
    t2s = np.arange(n, dtype=np.int32) if exclude_self else None
    if name == "lap3d":
        pot, grad = oracle.p2p_dense(3, tgt, src, e, t2s=t2s, want_grad=True)
        return [pot, grad[0], tgt[2]*pot]
    if name == "lap3d_dipole":
        pot, grad = oracle.p2p_dense(3, tgt, src, np.zeros(n), wd=[0*e, -e, 0*e], t2s=t2s,
                                     want_grad=True)
        return [pot, grad[2]]
    if name == "lap2d":
        pot, grad = oracle.p2p_dense(2, tgt, src, e, t2s=t2s, want_grad=True)
        return [pot, grad[1]]
    ec = e.astype(np.complex128)
    pot, grad = oracle.p2p_dense(13, tgt, src, ec, wd=[-ec, 0*ec, 0*ec], t2s=t2s,
                                 k=5.0 + 0.5j, want_grad=True)
    return [pot, grad[1]]


@pytest.mark.parametrize("name", ["lap3d", "lap3d_dipole", "lap2d", "helm3d"])
@pytest.mark.parametrize("exclude_self", [False, True])
def test_p2p_matrix(actx, name, exclude_self):
    from sumpy_amd import P2P, P2PMatrixGenerator, P2PMatrixSubsetGenerator
    dim, base, skn, tkn, kw = _kernels(name)
    rng = np.random.default_rng(4)
    n = 300
    src = rng.random((dim, n))
    tgt = src if exclude_self else rng.random((dim, 200)) + 0.05
    nt = tgt.shape[1]
    extra = dict(kw)
    if exclude_self:
        extra["target_to_source"] = actx.from_numpy(np.arange(n, dtype=np.int32))
    mats = P2PMatrixGenerator(tkn, exclude_self=exclude_self, source_kernels=skn)(
        actx, actx.from_numpy(tgt), actx.from_numpy(src), **extra)
    # every source kernel at unit strength: the matrix applied to a strength
    # vector is the direct sum with that strength for all of them
    q = rng.standard_normal(n)
    applied = P2P(tkn, exclude_self=exclude_self, source_kernels=skn)(
        actx, actx.from_numpy(tgt), actx.from_numpy(src), [actx.from_numpy(q) for _ in skn],
        **extra)
    for m, a in zip(mats, applied):
        m = actx.to_numpy(m)
        assert m.shape == (nt, n)
        assert rel(m @ q, actx.to_numpy(a)) < 1e-12
    if exclude_self:
        for m in mats:
            assert np.all(np.diag(actx.to_numpy(m)) == 0)
    # columns against the oracle's direct sum of a unit source, for every
    # target kernel of the case
    for j in (0, 17, n - 1):
        cols = _oracle_columns(name, tgt, src, j, exclude_self)
        for m, col in zip(mats, cols):
            assert rel(actx.to_numpy(m)[:, j], col) < 1e-13, (name, j)
    # subset entries
    ti = rng.integers(0, nt, 500).astype(np.int32)
    si = rng.integers(0, n, 500).astype(np.int32)
    sub = P2PMatrixSubsetGenerator(tkn, exclude_self=exclude_self, source_kernels=skn)(
        actx, actx.from_numpy(tgt), actx.from_numpy(src), tgtindices=actx.from_numpy(ti),
        srcindices=actx.from_numpy(si), **extra)
    for m, s in zip(mats, sub):
        assert np.array_equal(actx.to_numpy(s), actx.to_numpy(m)[ti, si])


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_layer_potential_matrix(actx, kind):
    from sumpy_amd import (
        LayerPotential,
        LayerPotentialMatrixGenerator,
        LayerPotentialMatrixSubsetGenerator,
    )
    from sumpy_amd.expansion.local import LineTaylorLocalExpansion
    from sumpy_amd.kernel import AxisTargetDerivative, DirectionalSourceDerivative, LaplaceKernel
    lap = LaplaceKernel(2)
    expn = LineTaylorLocalExpansion(lap, 6)
    t = np.linspace(0, 2*np.pi, 400, endpoint=False)
    r = 1 + 0.25*np.cos(5*t)
    nodes = np.stack([r*np.cos(t), r*np.sin(t)])
    tang = np.gradient(nodes, axis=1)
    normals = np.stack([tang[1], -tang[0]])/np.linalg.norm(tang, axis=0)
    centers = nodes - 0.02*normals
    radii = np.full(t.shape, 0.02)
    skn = [lap] if kind == "slp" else [DirectionalSourceDerivative(lap, "nrm")]
    tkn = [lap, AxisTargetDerivative(0, lap)]
    kw = {} if kind == "slp" else {"nrm": actx.from_numpy(normals)}
    args = (actx, actx.from_numpy(nodes), actx.from_numpy(nodes), actx.from_numpy(centers))
    mats = LayerPotentialMatrixGenerator(expn, source_kernels=skn, target_kernels=tkn)(
        *args, actx.from_numpy(radii), **kw)
    dens = np.sin(3*t)
    applied = LayerPotential(expn, source_kernels=skn, target_kernels=tkn)(
        *args, [actx.from_numpy(dens)], actx.from_numpy(radii), **kw)
    for m, a in zip(mats, applied):
        assert rel(actx.to_numpy(m) @ dens, actx.to_numpy(a)) < 1e-12
    # columns against the oracle's line-Taylor sum of a unit density
    srcmode = 0 if kind == "slp" else 1
    for j in (0, 123, 399):
        e = np.zeros(400)
        e[j] = 1
        for m, axis in zip(mats, (-1, 0)):
            col = oracle.qbx2d(6, srcmode, axis, nodes, centers, radii, nodes, normals, e)
            assert rel(actx.to_numpy(m)[:, j], col) < 1e-12, (kind, axis, j)
    rng = np.random.default_rng(2)
    ti = rng.integers(0, 400, 300).astype(np.int32)
    si = rng.integers(0, 400, 300).astype(np.int32)
    sub = LayerPotentialMatrixSubsetGenerator(expn, source_kernels=skn, target_kernels=tkn)(
        *args, actx.from_numpy(radii), actx.from_numpy(ti), actx.from_numpy(si), **kw)
    for m, s in zip(mats, sub):
        assert np.array_equal(actx.to_numpy(s), actx.to_numpy(m)[ti, si])
