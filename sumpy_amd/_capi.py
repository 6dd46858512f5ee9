"""ctypes binding of ``libsumpyhip.so`` (C-ABI in ``include/sumpyhip.h``).

The library is required: there is no CPU fallback.  Every entry point returns
an ``int`` status; a non-zero status raises :class:`SumpyHipError` carrying
``sumpyhip_last_error()``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsumpyhip.so")

SUMPYHIP_EINVAL = 1


class SumpyHipError(RuntimeError):
    pass


class ExpnDesc(ctypes.Structure):
    """Mirror of ``struct sumpyhip_expn``."""
    _fields_ = [
        ("order", ctypes.c_int), ("nfull", ctypes.c_int), ("nstored", ctypes.c_int),
        ("full_ids", ctypes.c_void_p), ("stored_ids", ctypes.c_void_p),
        ("stored_in_full", ctypes.c_void_p), ("cube_index", ctypes.c_void_p),
        ("proj_row_starts", ctypes.c_void_p), ("proj_cols", ctypes.c_void_p),
        ("proj_vals", ctypes.c_void_p), ("projT_col_starts", ctypes.c_void_p),
        ("projT_rows", ctypes.c_void_p), ("projT_vals", ctypes.c_void_p),
        ("horner_order", ctypes.c_void_p), ("proj_nnz", ctypes.c_int),
    ]


class M2LLevel(ctypes.Structure):
    """Mirror of ``struct sumpyhip_m2l_level``."""
    _fields_ = [
        ("ntiles", ctypes.c_int), ("nsplit", ctypes.c_int),
        ("tile_targets", ctypes.c_void_p), ("tile_slot_starts", ctypes.c_void_p),
        ("slot_class", ctypes.c_void_p), ("slot_sources", ctypes.c_void_p),
        ("class_data", ctypes.c_void_p), ("packed_src", ctypes.c_void_p),
        ("src_base_ibox", ctypes.c_int), ("work", ctypes.c_void_p),
        ("tgt_base_ibox", ctypes.c_int), ("partial", ctypes.c_void_p),
        ("tile_width", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_longlong

_SIGNATURES = {
    "sumpyhip_device_info": [_P, _P, _P],
    "sumpyhip_p2p_dense": [_I, _I, _D, _D, _D, _I, _P, _P, _P, _I, _P, _P, _P,
                           _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sumpyhip_p2p_csr": [_I, _I, _D, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P,
                         _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                         _P, _P, _P, _P, _P],
    "sumpyhip_p2m": [_I, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                     _D, _P, _I, _P],
    "sumpyhip_p2l": [_I, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P,
                     _P, _P, _D, _P, _I, _P],
    "sumpyhip_m2m": [_I, _P, _P, _I, _P, _P, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_l2l": [_I, _P, _P, _I, _P, _P, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_l2p": [_I, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P, _D, _D, _P, _I,
                     _P, _P, _P, _P, _P],
    "sumpyhip_m2p": [_I, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _D, _D,
                     _P, _I, _P, _P, _P, _P, _P],
    "sumpyhip_h2d_p2m": [_I, _I, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _D, _P, _I, _P],
    "sumpyhip_h2d_p2l": [_I, _I, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _D, _P, _I,
                         _P],
    "sumpyhip_h2d_m2m": [_I, _I, _I, _D, _D, _I, _P, _P, _I, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_h2d_l2l": [_I, _I, _I, _D, _D, _I, _P, _P, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_h2d_m2l_data": [_I, _I, _I, _D, _D, _I, _P, _I, _I, _P, _P],
    "sumpyhip_h2d_m2l": [_I, _I, _I, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P, _I, _D, _D, _P, _I, _P,
                         _I, _P],
    "sumpyhip_h2d_l2p": [_I, _I, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P, _D, _P, _I, _P, _P, _P, _P],
    "sumpyhip_h2d_m2p": [_I, _I, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _D, _P, _I, _P, _P,
                         _P, _P],
    "sumpyhip_h2d_p2p_csr": [_I, _D, _D, _I] + [_P]*19,
    "sumpyhip_h2d_p2p_dense": [_I, _D, _D, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                               _P],
    "sumpytree_gpu_morton": [_I, _L, _P, _P, _P, _P, _D, _I, _P, _P],
    "sumpytree_gpu_children": [_I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P],
    "sumpytree_gpu_geometry": [_I, _I, _P, _P, _P, _D, _P, _P, _P],
    "sumpytree_gpu_colleagues": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sumpytree_gpu_list": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P,
                           _P, _P, _P, _P, _P],
    "sumpytree_gpu_list2_classes": [_I, _P, _P, _P, _I, _P, _P],
    "sumpyhip_p2m_levels": [_I, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                            _I, _P, _P, _I, _P],
    "sumpyhip_l2p_levels": [_I, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P, _D, _P,
                            _P, _P, _P, _P, _I, _P],
    "sumpyhip_m2p_levels": [_I, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P,
                            _D, _P, _P, _P, _P, _P, _I, _P],
    "sumpyhip_p2l_levels": [_I, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P,
                            _P, _P, _I, _P, _P, _I, _P],
    "sumpyhip_m2p_helmholtz": [_P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _D,
                               _D, _P, _I, _P, _P, _P, _P, _I, _P],
    "sumpyhip_p2l_helmholtz": [_P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P,
                               _P, _P, _P, _D, _P, _I, _I, _P],
    "sumpyhip_m2l_csr_helmholtz": [_I, _I, _P, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _I, _D,
                                   _D, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_m2l_class_data_helmholtz": [_I, _I, _P, _I, _P, _I, _I, _D, _D, _D, _P, _P],
    "sumpyhip_octant_translate": [_I, _P, _P, _P, _I, _P, _I, _I, _P, _P, _I, _P, _I, _P,
                                  _P],
    "sumpyhip_qbx2d_adjoint": [_I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _D, _P,
                               _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sumpyhip_m2l_pack_bf16_class_data": [_I, _P, _I, _I, _I, _P, _P, _P],
    "sumpyhip_m2l_class_data": [_I, _I, _I, _P, _I, _P, _I, _I, _D, _P, _P],
    "sumpyhip_m2l_preprocess": [_I, _I, _I, _P, _I, _P, _P, _P],
    "sumpyhip_m2l_apply": [_I, _I, _P, _I, _I, _I, _I, _P, _P, _I, _P],
    "sumpyhip_m2l_apply_noalloc": [_I, _I, _P, _I, _I, _I, _I, _P, _P, _I, _P],
    "sumpyhip_m2l_postprocess": [_I, _I, _I, _I, _P, _P, _P, _D, _P, _P],
    "sumpyhip_m2l_csr": [_I, _I, _I, _P, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P,
                         _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_p2p_matrix": [_I, _D, _D, _D, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                            _P, _L, _P, _P, _P, _P, _P, _P, _P],
    "sumpyhip_qbx2d_matrix": [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _L, _P, _P,
                              _D, _I, _P, _P],
    "sumpyhip_axpy_weights": [_I, _I, _D, _P, _P, _P, _P],
    "sumpyhip_combine_outputs": [_I, _I, _I, _P, _P, _P, _P, _I, _P],
    "sumpyhip_t2d_p2m": [_I, _P, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _D, _P, _I,
                         _P],
    "sumpyhip_t2d_l2p": [_I, _P, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P, _D, _P, _I, _D, _D, _P,
                         _P, _P, _P],
    "sumpyhip_t2d_m2p": [_I, _P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _D, _P, _I, _D,
                         _D, _P, _P, _P, _P],
    "sumpyhip_t2d_p2l": [_I, _P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _D,
                         _P, _I, _P],
    "sumpyhip_t2d_m2m": [_I, _P, _P, _D, _D, _I, _P, _P, _I, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_t2d_l2l": [_I, _P, _P, _D, _D, _I, _P, _P, _P, _I, _D, _D, _P, _I, _P, _I, _P],
    "sumpyhip_t2d_m2l_data": [_I, _I, _D, _D, _I, _P, _I, _I, _D, _P, _P],
    "sumpyhip_t2d_m2l": [_I, _P, _P, _D, _D, _I, _P, _P, _P, _P, _I, _P, _P, _I, _D, _D, _P, _I,
                         _P, _I, _P],
    "sumpyhip_qbx_line": [_I, _I, _I, _I, _I, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P,
                          _P, _P, _P, _P, _P, _D, _D, _I, _P, _P],
    "sumpyhip_qbx_line_matrix": [_I, _I, _I, _I, _I, _D, _D, _I, _P, _P, _P, _P, _P, _P, _I,
                                 _P, _P, _P, _P, _P, _P, _L, _P, _P, _D, _D, _I, _P, _P],
    "sumpyhip_qbx2d": [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P,
                       _D, _P, _P],
}

EXPORTED = ["sumpyhip_last_error", *_SIGNATURES]

_lib = None


def load(path: str = LIB_PATH):
    """Load the library (no GPU needed) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SumpyHipError(
            f"{path} is missing: the HIP extension must be built "
            "(__graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    lib.sumpyhip_last_error.restype = ctypes.c_char_p
    lib.sumpyhip_last_error.argtypes = []
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    lib = load()
    # expansion descriptors are passed by address (struct sumpyhip_expn *)
    args = [ctypes.addressof(a) if isinstance(a, ExpnDesc) else a for a in args]
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.sumpyhip_last_error().decode(errors="replace")
        if rc == SUMPYHIP_EINVAL:
            raise ValueError(f"{name}: {msg}")
        raise SumpyHipError(f"{name} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    """Device pointer of a tensor (None passes NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"expected a torch.Tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def require_gpu(t: torch.Tensor) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise SumpyHipError(
            "sumpy_amd operators run on the GPU only: pass device tensors "
            "(there is no CPU fallback)")


def stream_of(actx) -> int:
    return actx.stream_handle


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))(*ptrs)
    return arr


def f64_array(vals):
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])


def np_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
