"""Multipole expansions (reference ``sumpy/expansion/multipole.py:70-419``).

Coefficients are ``c_mi = sum_s q_s (a/rho)^mi / mi!`` (``a`` = centre -
source) compressed by ``P^T`` for the PDE-conforming form; evaluation is
``sum_mi c_mi rho^|mi| d^mi G(b)``.  Formation, translation and evaluation run
in the HIP kernels of ``sumpy_amd/csrc/expansions.hip``.
"""

from __future__ import annotations

from dataclasses import dataclass

from sumpy_amd.expansion import ExpansionBase, VolumeTaylorExpansionMixin


class MultipoleExpansionBase(ExpansionBase):
    pass


@dataclass(frozen=True)
class VolumeTaylorMultipoleExpansion(VolumeTaylorExpansionMixin, MultipoleExpansionBase):
    """Full Taylor multipole (``multipole.py:407-410``)."""

    compressed = False


@dataclass(frozen=True)
class LinearPDEConformingVolumeTaylorMultipoleExpansion(
        VolumeTaylorExpansionMixin, MultipoleExpansionBase):
    """PDE-compressed Taylor multipole (``multipole.py:413-416``)."""

    compressed = True


@dataclass(frozen=True)
class HankelBased2DMultipoleExpansion(MultipoleExpansionBase):
    """2D Hankel multipole (``multipole.py:424-497``): coefficients
    ``c = -p..p`` at storage index ``c + p``; the operators run the HIP
    kernels of ``csrc/h2d.hip`` (:mod:`sumpy_amd.h2d`)."""

    def get_coefficient_identifiers(self):
        return [(i,) for i in range(-self.order, self.order + 1)]

    def get_storage_index(self, mi):
        ind, = mi
        return self.order + ind


@dataclass(frozen=True)
class H2DMultipoleExpansion(HankelBased2DMultipoleExpansion):
    """For the 2D Helmholtz kernel (``multipole.py:532-550``)."""

    def __post_init__(self):
        from sumpy_amd.kernel import HelmholtzKernel
        base = self.kernel.get_base_kernel()
        if not (isinstance(base, HelmholtzKernel) and base.dim == 2):
            raise TypeError(f"{type(self).__name__} can only be applied to 2D "
                            f"HelmholtzKernel: {self.kernel!r}")


@dataclass(frozen=True)
class Y2DMultipoleExpansion(HankelBased2DMultipoleExpansion):
    """For the 2D Yukawa kernel (``multipole.py:551-567``): the Hankel
    expansion with the Bessel argument scaled by ``i lam``."""

    def __post_init__(self):
        from sumpy_amd.kernel import YukawaKernel
        base = self.kernel.get_base_kernel()
        if not (isinstance(base, YukawaKernel) and base.dim == 2):
            raise TypeError(f"{type(self).__name__} can only be applied to 2D "
                            f"YukawaKernel: {self.kernel!r}")


__all__ = [
    "H2DMultipoleExpansion",
    "Y2DMultipoleExpansion",
    "HankelBased2DMultipoleExpansion",
    "LinearPDEConformingVolumeTaylorMultipoleExpansion",
    "MultipoleExpansionBase",
    "VolumeTaylorMultipoleExpansion",
]
