"""Kernel descriptors with the reference's names and wrapper algebra.

The reference builds symbolic expressions for each kernel
(``sumpy/kernel.py:639-691`` Laplace, ``:754-844`` Helmholtz) and generates
code from them.  Here a kernel is a small immutable descriptor: the HIP
kernels implement the Green's functions by hand, and the wrappers
(``AxisSourceDerivative`` ``kernel.py:2288-2327``, ``AxisTargetDerivative``
``:2330-2385``, ``DirectionalSourceDerivative`` ``:2457-2507``,
``TargetPointMultiplier`` ``:2510-2588``) are reduced to two linear maps:

* on the source side every source kernel becomes a *charge* weight and a
  *dipole* vector (``pot = w0 G(d) + w . grad_d G(d)``, ``d = target - source``);
* on the target side every target kernel becomes a combination of the
  potential, its target gradient and the target coordinates.

That covers every first-derivative combination the reference exercises in its
FMM tests (``test_fmm.py:386-741``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


class ScalarKernel:
    """Base of all kernel descriptors (reference ``kernel.py:260-469``)."""

    dim: int
    is_translation_invariant = True

    def get_base_kernel(self) -> ScalarKernel:
        return self

    @property
    def is_complex_valued(self) -> bool:
        return self.get_base_kernel().is_complex_valued

    def get_global_scaling_const(self) -> float:
        return self.get_base_kernel().get_global_scaling_const()

    def get_args(self):
        return self.get_base_kernel().get_args()

    def get_source_args(self):
        return []


@dataclass(frozen=True)
class LaplaceKernel(ScalarKernel):
    """``G = 1/|d|`` (3D) or ``log|d|`` (2D) before the global scaling
    ``1/(4 pi)`` resp. ``-1/(2 pi)`` (reference ``kernel.py:648-658``)."""

    dim: int

    def __post_init__(self):
        if self.dim not in (2, 3):
            raise NotImplementedError(f"unsupported dimension: '{self.dim}'")

    @property
    def is_complex_valued(self) -> bool:
        return False

    def get_global_scaling_const(self) -> float:
        if self.dim == 3:
            return 1.0 / (4.0 * math.pi)
        return 1.0 / (-2.0 * math.pi)

    def get_args(self):
        return []

    def __str__(self):
        return f"LapKnl{self.dim}D"


@dataclass(frozen=True)
class HelmholtzKernel(ScalarKernel):
    """``G = exp(i k |d|)/|d|`` (3D) with global scaling ``1/(4 pi)``, or
    ``G = H0^(1)(k |d|)`` (2D) with scaling ``i/4`` (reference
    ``kernel.py:754-844``).  ``k`` is passed at call time under
    :attr:`helmholtz_k_name`; it is real unless *allow_evanescent*.  The 2D
    kernel runs with the H2D expansions and its own P2P
    (:mod:`sumpy_amd.h2d`)."""

    dim: int
    helmholtz_k_name: str = "k"
    allow_evanescent: bool = False

    def __post_init__(self):
        if self.dim not in (2, 3):
            raise NotImplementedError(f"unsupported dimension: '{self.dim}'")

    @property
    def is_complex_valued(self) -> bool:
        return True

    def get_global_scaling_const(self):
        return 1.0 / (4.0 * math.pi) if self.dim == 3 else 0.25j

    def get_args(self):
        return [self.helmholtz_k_name]

    def __str__(self):
        return f"HelmKnl{self.dim}D({self.helmholtz_k_name})"


@dataclass(frozen=True)
class YukawaKernel(ScalarKernel):
    """``G = (i/4) H0^(1)(i lam |d|) = K0(lam |d|)/(2 pi)`` (2D) or
    ``G = exp(-lam |d|)/(4 pi |d|)`` (3D) (reference ``kernel.py:848-923``,
    which writes the 2D kernel through the Hankel function with the scaling
    ``i/4``).  ``lam`` is passed at call time under
    :attr:`yukawa_lambda_name`.  Complex-valued as in the reference.  The 2D
    kernel runs with the Y2D expansions and its own P2P (:mod:`sumpy_amd.h2d`);
    the 3D kernel is the 3D Helmholtz kernel at k = i lam (P2P, the full-Taylor
    FMM) and has its own line-Taylor layer potential (:mod:`sumpy_amd.qbx`)."""

    dim: int
    yukawa_lambda_name: str = "lam"

    def __post_init__(self):
        if self.dim not in (2, 3):
            raise NotImplementedError(f"unsupported dimension: '{self.dim}'")

    @property
    def is_complex_valued(self) -> bool:
        return True

    def get_global_scaling_const(self):
        return 0.25j if self.dim == 2 else 1.0 / (4.0 * math.pi)

    def get_args(self):
        return [self.yukawa_lambda_name]

    def __str__(self):
        return f"YukKnl{self.dim}D({self.yukawa_lambda_name})"


class KernelWrapper(ScalarKernel):
    inner_kernel: ScalarKernel

    @property
    def dim(self):
        return self.inner_kernel.dim

    def get_base_kernel(self):
        return self.inner_kernel.get_base_kernel()

    def get_source_args(self):
        return self.inner_kernel.get_source_args()


@dataclass(frozen=True)
class AxisSourceDerivative(KernelWrapper):
    """``d/dy_axis G(x - y)`` (reference ``kernel.py:2288-2327``)."""

    axis: int
    inner_kernel: ScalarKernel

    def __str__(self):
        return f"d/dy{self.axis} {self.inner_kernel}"


@dataclass(frozen=True)
class AxisTargetDerivative(KernelWrapper):
    """``d/dx_axis`` of the inner kernel (reference ``kernel.py:2330-2385``)."""

    axis: int
    inner_kernel: ScalarKernel

    def __str__(self):
        return f"d/dx{self.axis} {self.inner_kernel}"


@dataclass(frozen=True)
class DirectionalSourceDerivative(KernelWrapper):
    """``n(y) . grad_y G(x - y)`` with ``n`` passed per source under
    *dir_vec_name* (reference ``kernel.py:2457-2507``; default name
    ``src_derivative_dir`` at ``kernel.py:2435-2436``)."""

    inner_kernel: ScalarKernel
    dir_vec_name: str = "src_derivative_dir"

    def get_source_args(self):
        return [self.dir_vec_name, *self.inner_kernel.get_source_args()]

    def __str__(self):
        return f"{self.dir_vec_name}.grad_y {self.inner_kernel}"


@dataclass(frozen=True)
class TargetPointMultiplier(KernelWrapper):
    """``x_axis G(x, y)`` (reference ``kernel.py:2510-2588``)."""

    axis: int
    inner_kernel: ScalarKernel

    def __str__(self):
        return f"x{self.axis} {self.inner_kernel}"


# {{{ transformation removers (reference kernel.py:2593-2760)

def remove_source_transforms(knl: ScalarKernel) -> ScalarKernel:
    """Strip source-side wrappers, keeping target-side ones."""
    if isinstance(knl, (AxisSourceDerivative, DirectionalSourceDerivative)):
        return remove_source_transforms(knl.inner_kernel)
    if isinstance(knl, AxisTargetDerivative):
        return AxisTargetDerivative(knl.axis,
                                    remove_source_transforms(knl.inner_kernel))
    if isinstance(knl, TargetPointMultiplier):
        return TargetPointMultiplier(knl.axis,
                                     remove_source_transforms(knl.inner_kernel))
    return knl


def remove_target_transforms(knl: ScalarKernel) -> ScalarKernel:
    """Strip target-side wrappers, keeping source-side ones."""
    if isinstance(knl, (AxisTargetDerivative, TargetPointMultiplier)):
        return remove_target_transforms(knl.inner_kernel)
    if isinstance(knl, AxisSourceDerivative):
        return AxisSourceDerivative(knl.axis,
                                    remove_target_transforms(knl.inner_kernel))
    if isinstance(knl, DirectionalSourceDerivative):
        return DirectionalSourceDerivative(
            remove_target_transforms(knl.inner_kernel), knl.dir_vec_name)
    return knl


SourceTransformationRemover = lambda: remove_source_transforms  # noqa: E731
TargetTransformationRemover = lambda: remove_target_transforms  # noqa: E731

# }}}


# {{{ source side: charge/dipole decomposition

@dataclass(frozen=True)
class SourceTerm:
    """``coeff * (strength) * [G | d/dd_axis G]`` contribution of one source
    kernel.  ``axis = -1`` is the charge term; ``dir_vec`` names a per-source
    direction array whose component ``axis`` multiplies the term."""

    axis: int
    coeff: float
    dir_vec: str | None = None


def source_terms(knl: ScalarKernel) -> list[SourceTerm]:
    """Decompose a source kernel into charge and dipole terms.

    ``d/dy_a G(x - y) = -dG/dd_a`` with ``d = x - y``
    (reference ``kernel.py:2308-2318``); ``n . grad_y`` gives ``-n_a dG/dd_a``
    (``kernel.py:2474-2490``).  Only first source derivatives are supported.
    """
    if isinstance(knl, (LaplaceKernel, HelmholtzKernel, YukawaKernel)):
        return [SourceTerm(-1, 1.0)]
    if isinstance(knl, AxisSourceDerivative):
        inner = knl.inner_kernel
        if not isinstance(inner, (LaplaceKernel, HelmholtzKernel, YukawaKernel)):
            raise NotImplementedError(
                "only first-order source derivatives are implemented")
        return [SourceTerm(knl.axis, -1.0)]
    if isinstance(knl, DirectionalSourceDerivative):
        inner = knl.inner_kernel
        if not isinstance(inner, (LaplaceKernel, HelmholtzKernel, YukawaKernel)):
            raise NotImplementedError(
                "only first-order source derivatives are implemented")
        return [SourceTerm(a, -1.0, knl.dir_vec_name) for a in range(knl.dim)]
    raise NotImplementedError(f"unsupported source kernel: {knl}")


def needs_dipole(source_kernels) -> bool:
    return any(t.axis >= 0 for k in source_kernels for t in source_terms(k))

# }}}


# {{{ target side: potential/gradient/point-multiplier decomposition

@dataclass(frozen=True)
class TargetTerm:
    """``coeff * x_{mult_axis} * out[deriv]`` where ``deriv = -1`` is the
    potential, ``0..dim-1`` a gradient component, and ``mult_axis = -1``
    means no coordinate factor."""

    deriv: int
    coeff: float
    mult_axis: int = -1


def target_terms(knl: ScalarKernel) -> list[TargetTerm]:
    """Linear decomposition of a target kernel.

    ``d/dx_a (x_j u) = delta_aj u + x_j du/dx_a`` (reference
    ``kernel.py:2548-2570`` with ``AxisTargetDerivative`` applied outside).
    """
    if isinstance(knl, (LaplaceKernel, HelmholtzKernel, YukawaKernel)):
        return [TargetTerm(-1, 1.0)]
    if isinstance(knl, TargetPointMultiplier):
        inner = target_terms(knl.inner_kernel)
        out = []
        for t in inner:
            if t.mult_axis >= 0:
                raise NotImplementedError("nested point multipliers")
            out.append(TargetTerm(t.deriv, t.coeff, knl.axis))
        return out
    if isinstance(knl, AxisTargetDerivative):
        inner = target_terms(knl.inner_kernel)
        out = []
        for t in inner:
            if t.deriv >= 0:
                raise NotImplementedError(
                    "only first-order target derivatives are implemented")
            out.append(TargetTerm(knl.axis, t.coeff, t.mult_axis))
            if t.mult_axis == knl.axis:
                out.append(TargetTerm(-1, t.coeff, -1))
        return out
    raise NotImplementedError(f"unsupported target kernel: {knl}")


def needs_gradient(target_kernels) -> bool:
    return any(t.deriv >= 0 for k in target_kernels for t in target_terms(k))

# }}}


def kernel_id(knl: ScalarKernel) -> int:
    """Integer id of the base kernel used across the C-ABI."""
    base = knl.get_base_kernel()
    if isinstance(base, LaplaceKernel):
        return 3 if base.dim == 3 else 2
    if isinstance(base, HelmholtzKernel):
        return 13 if base.dim == 3 else 12
    if isinstance(base, YukawaKernel) and base.dim == 2:
        return 22
    if isinstance(base, YukawaKernel) and base.dim == 3:
        # e^{-lam r}/(4 pi r): the 3D Helmholtz kernel at k = i lam
        return 13
    raise NotImplementedError(f"unsupported kernel: {knl}")


def value_dtype(knl: ScalarKernel):
    return np.dtype(np.complex128 if knl.is_complex_valued else np.float64)
