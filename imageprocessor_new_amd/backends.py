"""Backend resolution and device plumbing (reference subpx/backends.py:28-59).

This package implements the reference's ``backend="gpu"`` path on MI355X: device
buffers are torch tensors on a HIP device, kernels come from libsubpx_hip.so. There is
no CPU implementation here (the reference's CPU path is OpenCV/NumPy code of its own);
``backend="cpu"`` raises NotImplementedError and a GPU request without a device raises
RuntimeError, as the reference does without CuPy (subpx/backends.py:31-32).
"""

from __future__ import annotations

from contextlib import contextmanager

import numpy as np
import torch

from . import _lib


def is_gpu_available() -> bool:
    return bool(torch.cuda.is_available()) and _lib.LIB_PATH.exists()


def resolve_backend(backend: str = "auto") -> str:
    if backend not in ("auto", "gpu", "cpu"):
        raise ValueError(f"backend must be 'auto', 'cpu' or 'gpu', got {backend!r}")
    if backend == "cpu":
        raise NotImplementedError("imageprocessor_new_amd implements the GPU backend only")
    if not torch.cuda.is_available():
        raise RuntimeError("backend='gpu' requested, but no HIP device is available.")
    _lib.load()
    return "gpu"


_initialised: set[int] = set()


@contextmanager
def gpu_device(device: int = 0):
    """Make ``device`` current for torch and the library for the duration."""
    device = int(device)
    with torch.cuda.device(device):
        if device not in _initialised:
            _lib.call("spx_init", device)
            _initialised.add(device)
        yield torch.device("cuda", device)


def stream_handle(stream: torch.cuda.Stream | None = None) -> int:
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream)


def to_device_u16(image, device: torch.device) -> torch.Tensor:
    """A contiguous (H, W) uint16 tensor on ``device`` (copies NumPy / host input)."""
    if isinstance(image, torch.Tensor):
        t = image
        if t.dtype != torch.uint16:
            if t.dtype in (torch.int16,):
                t = t.view(torch.uint16)
            else:
                t = t.to(torch.int32).clamp_(0, 65535).to(torch.uint16)
        return t.to(device).contiguous()
    a = np.asarray(image)
    if a.ndim != 2:
        raise ValueError("Expected 2D grayscale image.")
    if a.dtype != np.uint16:
        if a.dtype.kind == "f":
            a = np.clip(np.rint(a), 0, 65535)
        a = a.astype(np.uint16)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def to_numpy(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)
