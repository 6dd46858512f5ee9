"""Host-side wavenumber rules (no GPU): a complex Helmholtz k needs
``allow_evanescent`` (as the reference's ``HelmholtzKernel`` does,
``sumpy/kernel.py:764-778``); the 2D Bessel and volume-Taylor expansions take
Im k > 0; the PDE-compressed 3D expansions refuse a complex k while the
full-Taylor ones take it; the 3D Yukawa kernel is Helmholtz at k = i lam."""

import pytest

from sumpy_amd import h2d, taylor2d
from sumpy_amd.expansion.local import (
    LinearPDEConformingVolumeTaylorLocalExpansion,
    VolumeTaylorLocalExpansion,
)
from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel, YukawaKernel, kernel_id
from sumpy_amd.tools import expansion_k, helmholtz_k, kernel_k


def test_2d_bessel_wavenumbers():
    assert h2d.wavenumber(HelmholtzKernel(2), {"k": 1.5}) == (1.5, 0.0)
    assert h2d.wavenumber(HelmholtzKernel(2, allow_evanescent=True), {"k": 1 + 2j}) == (1.0, 2.0)
    with pytest.raises(ValueError, match="allow_evanescent"):
        h2d.wavenumber(HelmholtzKernel(2), {"k": 1 + 2j})
    with pytest.raises(NotImplementedError):
        h2d.wavenumber(HelmholtzKernel(2, allow_evanescent=True), {"k": 1 - 2j})
    assert h2d.wavenumber(YukawaKernel(2), {"lam": 3.0}) == (3.0, 0.0)


def test_2d_taylor_wavenumbers():
    ok = VolumeTaylorLocalExpansion(HelmholtzKernel(2, allow_evanescent=True), 4)
    assert taylor2d._k(ok, {"k": 0.5 + 0.25j}) == (0.5, 0.25)
    with pytest.raises(ValueError, match="allow_evanescent"):
        taylor2d._k(VolumeTaylorLocalExpansion(HelmholtzKernel(2), 4), {"k": 0.5 + 0.25j})
    assert taylor2d._k(VolumeTaylorLocalExpansion(LaplaceKernel(2), 4), {}) == (0.0, 0.0)


def test_3d_expansion_wavenumbers():
    ev = HelmholtzKernel(3, allow_evanescent=True)
    full = VolumeTaylorLocalExpansion(ev, 4)
    comp = LinearPDEConformingVolumeTaylorLocalExpansion(ev, 4)
    kw = {"k": 2 + 0.5j}
    assert expansion_k(ev, kw, full) == 0.0
    assert helmholtz_k(ev, kw) == (2.0, 0.5)
    with pytest.raises(NotImplementedError, match="full-Taylor"):
        expansion_k(ev, kw, comp)
    with pytest.raises(ValueError, match="allow_evanescent"):
        helmholtz_k(HelmholtzKernel(3), kw)
    assert expansion_k(HelmholtzKernel(3), {"k": 3.0}, comp) == 3.0


def test_3d_yukawa_is_helmholtz_at_imaginary_k():
    yk = YukawaKernel(3)
    assert kernel_id(yk) == kernel_id(HelmholtzKernel(3))
    assert kernel_k(yk, {"lam": 2.0}) == 2j
    assert helmholtz_k(yk, {"lam": 2.0}) == (0.0, 2.0)
    assert kernel_k(YukawaKernel(2), {"lam": 2.0}) == 2.0
