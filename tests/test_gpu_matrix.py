"""Matrix generators (reference ``sumpy/p2p.py:292-477``, ``sumpy/qbx.py:
342-539``): the full matrix times a strength vector equals the direct sum of
the applier (P2P / LayerPotential, themselves pinned to the oracle); single
columns equal the oracle's direct sum for a unit source; the subset generator
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
    # one column against the oracle (plain potential, charge sources)
    if name in ("lap3d", "lap2d"):
        j = 17
        w = np.zeros(n)
        w[j] = 1
        knl = 3 if dim == 3 else 2
        col, _ = oracle.p2p_dense(knl, tgt, src, w,
                                  t2s=np.arange(n) if exclude_self else None)
        assert rel(actx.to_numpy(mats[0])[:, j], col) < 1e-13
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
    rng = np.random.default_rng(2)
    ti = rng.integers(0, 400, 300).astype(np.int32)
    si = rng.integers(0, 400, 300).astype(np.int32)
    sub = LayerPotentialMatrixSubsetGenerator(expn, source_kernels=skn, target_kernels=tkn)(
        *args, actx.from_numpy(radii), actx.from_numpy(ti), actx.from_numpy(si), **kw)
    for m, s in zip(mats, sub):
        assert np.array_equal(actx.to_numpy(s), actx.to_numpy(m)[ti, si])
