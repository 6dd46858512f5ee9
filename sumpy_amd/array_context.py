"""Array context over PyTorch-ROCm tensors.

Stands where the reference's ``PyOpenCLArrayContext`` stands
(``sumpy/array_context.py:62-157``): every operator receives it as its first
argument.  Device arrays are ``torch.Tensor`` objects on the context's device;
all HIP kernels are enqueued on the context's current HIP stream (the
reference's single in-order command queue).
"""

from __future__ import annotations

import numpy as np
import torch


_NP_TO_TORCH = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.complex64): torch.complex64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.int16): torch.int16,
    np.dtype(np.uint16): torch.int16,
    np.dtype(np.int8): torch.int8,
    np.dtype(np.bool_): torch.bool,
}


def torch_dtype(dtype) -> torch.dtype:
    if isinstance(dtype, torch.dtype):
        return dtype
    return _NP_TO_TORCH[np.dtype(dtype)]


def numpy_dtype(dtype: torch.dtype) -> np.dtype:
    for k, v in _NP_TO_TORCH.items():
        if v == dtype:
            return k
    raise KeyError(dtype)


class _NumpyNamespace:
    def __init__(self, actx):
        self._actx = actx

    def zeros(self, shape, dtype=np.float64):
        return torch.zeros(shape, dtype=torch_dtype(dtype), device=self._actx.device)

    def zeros_like(self, ary):
        return torch.zeros_like(ary)

    def max(self, ary):
        return torch.max(ary)

    def abs(self, ary):
        return torch.abs(ary)


class TorchArrayContext:
    """Holds the device and the stream that every operator launches on."""

    def __init__(self, device: str | torch.device = "cuda"):
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.np = _NumpyNamespace(self)

    @property
    def stream_handle(self) -> int:
        """Raw ``hipStream_t`` of the current stream on this device."""
        return torch.cuda.current_stream(self.device).cuda_stream

    def from_numpy(self, ary):
        if isinstance(ary, np.ndarray) and ary.dtype.char == "O":
            return np.array([self.from_numpy(a) for a in ary] + [None],
                            dtype=object)[:-1]
        return torch.from_numpy(np.ascontiguousarray(ary)).to(self.device)

    def to_numpy(self, ary):
        if isinstance(ary, torch.Tensor):
            return ary.detach().cpu().numpy()
        if isinstance(ary, np.ndarray) and ary.dtype.char == "O":
            return np.array([self.to_numpy(a) for a in ary] + [None],
                            dtype=object)[:-1]
        if isinstance(ary, (list, tuple)):
            return type(ary)(self.to_numpy(a) for a in ary)
        return np.asarray(ary)

    def freeze(self, ary):
        return ary

    def thaw(self, ary):
        return ary

    def synchronize(self):
        torch.cuda.synchronize(self.device)


def obj_array(items) -> np.ndarray:
    """1-D numpy object array holding the given arrays (the reference's
    ``pytools.obj_array.new_1d``)."""
    items = list(items)
    out = np.empty(len(items), dtype=object)
    for i, a in enumerate(items):
        out[i] = a
    return out
