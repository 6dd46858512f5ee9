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


class FourierBesselM2LTranslation(M2LTranslationBase):
    """M2L of the 2D Bessel locals (reference ``m2l.py:1008-1115``): the
    class data are the Hankel row ``H_n(k|d|) e^{i n th_d}``,
    ``n = -(ps+pt)..(ps+pt)`` (``2 (ps + pt) + 1`` complex values, stored
    as FP64 (re, im) pairs); the multipole pre-scaling ``rs^|m|`` and the
    local post-scaling ``(-1)^j rt^|j|`` are folded into the contraction
    kernel, so no separate pre/postprocessing arrays are used."""

    use_fft = False
    use_preprocessing = False
    precision = "fp64"

    def preprocessed_dtype(self, dtype):
        return np.dtype(dtype)

    def translation_classes_dependent_ndata(self, tgt_expansion, src_expansion):
        return 2*(2*tgt_expansion.order + 2*src_expansion.order + 1)

    def preprocess_multipole_nexprs(self, tgt_expansion, src_expansion):
        return 2*src_expansion.order + 1

    def postprocess_local_nexprs(self, tgt_expansion, src_expansion):
        return 2*tgt_expansion.order + 1


class VolumeTaylor2DM2LTranslation(M2LTranslationBase):
    """M2L of the 2D Volume-Taylor locals: the class data are the scaled
    derivatives ``rho^|m| d^m G(d_c)``, ``|m| <= ps + pt``, at index
    ``|m|(|m|+1)/2 + m_y`` (complex values as (re, im) pairs), the
    Toeplitz form ``L_t = sum_s M_s D[t + s]`` of ``m2l.py:363-389`` in 2D;
    no separate pre/postprocessing arrays (:mod:`sumpy_amd.taylor2d`)."""

    use_fft = False
    use_preprocessing = False
    precision = "fp64"

    def preprocessed_dtype(self, dtype):
        return np.dtype(dtype)

    def translation_classes_dependent_ndata(self, tgt_expansion, src_expansion):
        from sumpy_amd.taylor2d import class_ndata
        return class_ndata(tgt_expansion, src_expansion)

    def preprocess_multipole_nexprs(self, tgt_expansion, src_expansion):
        return len(src_expansion)

    def postprocess_local_nexprs(self, tgt_expansion, src_expansion):
        return len(tgt_expansion)


class FourierBesselM2LWithFFT(FourierBesselM2LTranslation):
    """Not supported, as in the reference (``m2l.py:1199-1215``)."""

    use_fft = True

    def __init__(self):
        raise ValueError("Bessel based expansions with FFT are not supported yet.")


def _is_fourier_bessel(local_expansion_class):
    from sumpy_amd.expansion.local import FourierBesselLocalExpansionMixin
    return isinstance(local_expansion_class, type) and \
        issubclass(local_expansion_class, FourierBesselLocalExpansionMixin)


class M2LTranslationClassFactoryBase:
    """Volume-Taylor locals: the dense class-matrix M2L; 2D Bessel locals:
    :class:`FourierBesselM2LTranslation` (reference ``m2l.py:100-173``)."""

    fourier_bessel = FourierBesselM2LTranslation

    def get_m2l_translation_class(self, base_kernel, local_expansion_class):
        if _is_fourier_bessel(local_expansion_class):
            return self.fourier_bessel
        if base_kernel.dim == 2:
            return VolumeTaylor2DM2LTranslation
        return VolumeTaylorM2LWithDenseClassMatrices


class FFTM2LTranslationClassFactory(M2LTranslationClassFactoryBase):
    fourier_bessel = FourierBesselM2LWithFFT


DefaultM2LTranslationClassFactory = M2LTranslationClassFactoryBase
NonFFTM2LTranslationClassFactory = M2LTranslationClassFactoryBase

__all__ = [
    "DefaultM2LTranslationClassFactory",
    "complex_rows",
    "FFTM2LTranslationClassFactory",
    "FourierBesselM2LTranslation",
    "FourierBesselM2LWithFFT",
    "M2LTranslationBase",
    "NonFFTM2LTranslationClassFactory",
    "VolumeTaylor2DM2LTranslation",
    "VolumeTaylorM2LWithDenseClassMatrices",
]
