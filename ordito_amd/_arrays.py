"""
Array plumbing between the caller's arrays and the library's device pointers.

Callers pass numpy arrays or torch tensors. Device memory and streams come from PyTorch-ROCm
(``torch.cuda`` is HIP here): inputs are staged as contiguous device tensors, results are float64
device tensors, returned as torch tensors on the device when the caller passed torch tensors and
as numpy arrays otherwise. Nothing here computes anything.
"""

from __future__ import annotations

from typing import Any

import numpy as np


def torch_module():
    import torch

    return torch


def require_gpu() -> None:
    torch = torch_module()
    if not torch.cuda.is_available():
        raise RuntimeError(
            "ordito_amd's heat path runs on an MI355X (gfx950) GPU through libheatgeo; no GPU is "
            "visible and there is no CPU fallback"
        )


def is_torch(value: Any) -> bool:
    try:
        torch = torch_module()
    except ImportError:  # pragma: no cover - torch is a dependency
        return False
    return isinstance(value, torch.Tensor)


def device_tensor(value: Any, dtype, *, shape_last: int | None = None):
    """``value`` as a contiguous tensor of ``dtype`` on the current GPU."""
    torch = torch_module()
    require_gpu()
    if is_torch(value):
        tensor = value.detach()
    else:
        tensor = torch.from_numpy(np.ascontiguousarray(np.asarray(value)))
    tensor = tensor.to(device="cuda", dtype=dtype).contiguous()
    if shape_last is not None:
        tensor = tensor.reshape(-1, shape_last)
    return tensor


def vertices_tensor(vertices: Any):
    torch = torch_module()
    t = device_tensor(vertices, torch.float32)
    if t.numel() % 3 != 0:
        raise ValueError(f"vertices must hold (n_vertices, 3) coordinates, got {tuple(t.shape)}")
    return t.reshape(-1, 3)


def faces_tensor(faces: Any):
    torch = torch_module()
    t = device_tensor(faces, torch.int32)
    if t.numel() % 3 != 0:
        raise ValueError(f"faces must hold 3 * n_faces indices, got {t.numel()}")
    return t.reshape(-1)


def sources_tensor(sources: Any):
    torch = torch_module()
    return device_tensor(sources, torch.int32).reshape(-1)


def like_input(result, template: Any):
    """Return ``result`` (a device tensor) as the caller's kind of array."""
    if is_torch(template):
        return result
    return result.cpu().numpy()


def pointer(tensor) -> int:
    return int(tensor.data_ptr()) if tensor is not None and tensor.numel() > 0 else 0


def current_stream() -> int:
    torch = torch_module()
    return int(torch.cuda.current_stream().cuda_stream)


_HIP = None


def device_copy(dst: int, src: int, nbytes: int) -> None:
    """Synchronous device-to-device copy through the HIP runtime (``hipMemcpy``)."""
    global _HIP
    import ctypes

    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipMemcpy.restype = ctypes.c_int
    status = _HIP.hipMemcpy(dst, src, nbytes, 4)  # hipMemcpyDefault
    if status != 0:
        raise RuntimeError(f"hipMemcpy failed with status {status}")
