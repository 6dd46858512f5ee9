"""M2L translation classes (reference ``sumpy/expansion/m2l.py``).

The reference offers a symbolic Toeplitz form (``VolumeTaylorM2LTranslation``,
``m2l.py:363-389``) and an FFT form (``VolumeTaylorM2LWithFFT``,
``m2l.py:913-1001``, the default).  This build has one translation class for
Volume-Taylor locals, :class:`VolumeTaylorM2LWithDenseClassMatrices`: the
class-dependent data keep the reference's non-FFT circulant layout, and the
translation of a level is a dense class-matrix contraction on MFMA.  The
factories of the reference return it for every request (an FFT M2L is not
built: the dense MFMA contraction replaces it).
"""

from __future__ import annotations

import numpy as np

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}


def _pad(n: int, m: int) -> int:
    return (n + m - 1)//m*m


class M2LTranslationBase:
    use_fft = False
    use_preprocessing = False

    def __eq__(self, other):
        return type(self) is type(other) and vars(self) == vars(other)

    def __hash__(self):
        return hash((type(self).__name__, tuple(sorted(vars(self).items()))))


class VolumeTaylorM2LWithDenseClassMatrices(M2LTranslationBase):
    """Class-batched dense M2L.

    * translation-class data: ``ncirc = prod(max_mi + 1)`` values per class,
      the circulant layout of ``m2l.py:403-537`` (FP64);
    * preprocessed multipoles: the stored coefficients packed into ``kpad``
      entries of the compute precision (``precision`` = ``"fp64"`` for the
      plain form, ``"fp32"`` for FP32 MFMA, ``"bf16x3"`` for split-BF16 MFMA:
      operands split as x = hi + lo in bf16, products hi*hi + hi*lo + lo*hi
      accumulated in FP32; FP32 rows padded to 32);
    * the translation writes ``npad`` entries of the compute precision per
      target box, which the postprocessing scales into FP64 locals.
    """

    use_fft = False
    use_preprocessing = True

    def __init__(self, precision: str = "fp64"):
        if precision not in PRECISIONS:
            raise ValueError(f"unknown M2L precision '{precision}'")
        self.precision = precision

    @property
    def precision_id(self) -> int:
        return PRECISIONS[self.precision]

    def preprocessed_dtype(self, dtype):
        """Real for every kernel: complex (Helmholtz) expansions are contracted
        through their real embedding (see :func:`complex_rows`)."""
        dtype = np.dtype(dtype)
        if dtype.kind == "c" and self.precision == "bf16x3":
            raise NotImplementedError("the split-BF16 M2L is implemented for real kernels")
        return np.dtype(np.float64 if self.precision == "fp64" else np.float32)

    def translation_classes_dependent_ndata(self, tgt_expansion, src_expansion):
        from sumpy_amd.expansion.terms import m2l_circulant_shape
        n = int(np.prod(m2l_circulant_shape(src_expansion.table, tgt_expansion.table)))
        return 2*n if tgt_expansion.is_complex_valued else n

    def preprocess_multipole_nexprs(self, tgt_expansion, src_expansion):
        rows = complex_rows(src_expansion)[0]
        return rows*_pad(len(src_expansion), 32 if self.precision == "bf16x3" else 4)

    def postprocess_local_nexprs(self, tgt_expansion, src_expansion):
        return complex_rows(tgt_expansion)[1]*_pad(len(tgt_expansion), 16)


def complex_rows(expansion):
    """(preprocessed rows, work rows) per box.

    A complex translation ``L = A M`` (Helmholtz) is contracted as the real
    block product ``[Lr; Li] = [Ar -Ai; Ai Ar] [Mr; Mi]``: each source box
    contributes the rows ``Mr, Mi, -Mi`` and each target box the rows
    ``Lr, Li``, and the class data hold ``Ar`` and ``Ai`` as two real classes.
    Real kernels use one row of each."""
    return (3, 2) if expansion.is_complex_valued else (1, 1)


class M2LTranslationClassFactoryBase:
    def get_m2l_translation_class(self, base_kernel, local_expansion_class):
        return VolumeTaylorM2LWithDenseClassMatrices


DefaultM2LTranslationClassFactory = M2LTranslationClassFactoryBase
NonFFTM2LTranslationClassFactory = M2LTranslationClassFactoryBase
FFTM2LTranslationClassFactory = M2LTranslationClassFactoryBase

__all__ = [
    "DefaultM2LTranslationClassFactory",
    "complex_rows",
    "FFTM2LTranslationClassFactory",
    "M2LTranslationBase",
    "NonFFTM2LTranslationClassFactory",
    "VolumeTaylorM2LWithDenseClassMatrices",
]
