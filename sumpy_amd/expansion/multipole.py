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


__all__ = [
    "LinearPDEConformingVolumeTaylorMultipoleExpansion",
    "MultipoleExpansionBase",
    "VolumeTaylorMultipoleExpansion",
]
