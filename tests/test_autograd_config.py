"""Configurations whose transpose the adjoint FMM pass does not compute are
refused (CPU: only wrangler attributes are inspected)."""

from functools import partial

import numpy as np
import pytest

from sumpy_amd.array_context import TorchArrayContext


def _wrangler(target_kernels=None, source_kernels=None, strength_usage=None,
              exclude_self=True, t2s=None, separate=False):
    from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
    from sumpy_amd.expansion.multipole import LinearPDEConformingVolumeTaylorMultipoleExpansion
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.kernel import LaplaceKernel
    from sumpy_amd.tree import build_tree_and_traversal
    actx = TorchArrayContext("cpu")
    rng = np.random.default_rng(0)
    src = rng.random((3, 500))
    tree, trav = build_tree_and_traversal(actx, src, rng.random((3, 200)) if separate else None,
                                          30)
    knl = LaplaceKernel(3)
    ti = SumpyTreeIndependentDataForWrangler(
        actx, partial(LinearPDEConformingVolumeTaylorMultipoleExpansion, knl),
        partial(LinearPDEConformingVolumeTaylorLocalExpansion, knl),
        target_kernels or [knl], exclude_self=exclude_self, source_kernels=source_kernels,
        strength_usage=strength_usage)
    if t2s is None and exclude_self:
        t2s = np.arange(tree.ntargets, dtype=np.int32)
    return SumpyExpansionWrangler(
        ti, trav, np.float64, fmm_level_to_order=lambda *a: 4,
        self_extra_kwargs={} if t2s is None else {"target_to_source": actx.from_numpy(t2s)})


def test_plain_configuration_is_accepted():
    from sumpy_amd.autograd import _check_adjoint_config
    _check_adjoint_config(_wrangler())
    _check_adjoint_config(_wrangler(exclude_self=False, separate=True))


@pytest.mark.parametrize("case", ["target_derivative", "two_targets", "dipole_source",
                                  "strength_usage", "permuted_t2s", "separate_exclude"])
def test_unsupported_configurations_raise(case):
    from sumpy_amd.autograd import _check_adjoint_config
    from sumpy_amd.kernel import AxisSourceDerivative, AxisTargetDerivative, LaplaceKernel
    knl = LaplaceKernel(3)
    kw = {
        "target_derivative": dict(target_kernels=[AxisTargetDerivative(0, knl)]),
        "two_targets": dict(target_kernels=[knl, AxisTargetDerivative(1, knl)]),
        "dipole_source": dict(source_kernels=[AxisSourceDerivative(2, knl)]),
        "strength_usage": dict(source_kernels=[knl, knl], strength_usage=[0, 1]),
        "permuted_t2s": dict(t2s=np.roll(np.arange(500, dtype=np.int32), 1)),
        "separate_exclude": dict(separate=True, t2s=np.arange(200, dtype=np.int32)),
    }[case]
    with pytest.raises(NotImplementedError):
        _check_adjoint_config(_wrangler(**kw))
