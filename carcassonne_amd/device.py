"""Device memory and stream plumbing (PyTorch-ROCm tensors as buffers; no torch compute)."""
import ctypes

import torch

_DEVICE = None


def device():
    global _DEVICE
    if _DEVICE is None:
        if not torch.cuda.is_available():
            raise RuntimeError("carcassonne_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        _DEVICE = torch.device("cuda", torch.cuda.current_device())
    return _DEVICE


def stream():
    """The caller's current HIP stream as a void* for the C ABI."""
    return ctypes.c_void_p(torch.cuda.current_stream(device()).cuda_stream)


def empty(shape, dtype=torch.float64):
    return torch.empty(tuple(int(s) for s in shape), dtype=dtype, device=device())


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t.numel() else ctypes.c_void_p(0)


def synchronize():
    torch.cuda.synchronize(device())
