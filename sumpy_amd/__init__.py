"""MI355X-native implementation of sumpy's FMM/QBX interaction path.

Exports the reference's operator names (``sumpy/__init__.py:52-66``); the
operators take a :class:`~sumpy_amd.array_context.TorchArrayContext` and
device tensors, and run hand-written gfx950 HIP kernels through the C-ABI of
``libsumpyhip.so`` (``include/sumpyhip.h``).
"""

from sumpy_amd.e2e import (
    E2EFromChildren,
    E2EFromCSR,
    E2EFromParent,
    M2LGenerateTranslationClassesDependentData,
    M2LPostprocessLocal,
    M2LPreprocessMultipole,
    M2LUsingTranslationClassesDependentData,
)
from sumpy_amd.e2p import E2PFromCSR, E2PFromSingleBox
from sumpy_amd.p2e import P2EFromCSR, P2EFromSingleBox
from sumpy_amd.p2p import P2P, P2PFromCSR, P2PMatrixGenerator, P2PMatrixSubsetGenerator
from sumpy_amd.qbx import (
    LayerPotential,
    LayerPotentialMatrixGenerator,
    LayerPotentialMatrixSubsetGenerator,
)

__all__ = [
    "P2P",
    "E2EFromCSR",
    "E2EFromChildren",
    "E2EFromParent",
    "E2PFromCSR",
    "E2PFromSingleBox",
    "M2LGenerateTranslationClassesDependentData",
    "M2LPostprocessLocal",
    "M2LPreprocessMultipole",
    "M2LUsingTranslationClassesDependentData",
    "P2EFromCSR",
    "P2EFromSingleBox",
    "P2PFromCSR",
    "P2PMatrixGenerator",
    "P2PMatrixSubsetGenerator",
    "LayerPotential",
    "LayerPotentialMatrixGenerator",
    "LayerPotentialMatrixSubsetGenerator",
]
