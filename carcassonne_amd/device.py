"""Device memory and stream plumbing (PyTorch-ROCm tensors as buffers; no torch compute)."""
import ctypes

import torch

_DEVICE = None


_ONES = None
ONES_LEN = 128


def device():
    global _DEVICE, _ONES
    if _DEVICE is None:
        if not torch.cuda.is_available():
            raise RuntimeError("carcassonne_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        d = torch.device("cuda", torch.cuda.current_device())
        # a constant vector of ones, complete before any stream can read it (products that add slices)
        _ONES = torch.ones(ONES_LEN, dtype=torch.float64, device=d)
        torch.cuda.synchronize(d)
        _DEVICE = d
    return _DEVICE


def ones():
    """ONES_LEN ones on the device (read only)."""
    device()
    return _ONES


_DEVICE_INDEX = None


def stream():
    """The caller's current HIP stream as a void* for the C ABI (the raw handle: this is called before
    every kernel launch, and building a torch Stream object each time costs microseconds)."""
    global _DEVICE_INDEX
    if _DEVICE_INDEX is None:
        _DEVICE_INDEX = device().index
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(_DEVICE_INDEX))


def empty(shape, dtype=torch.float64):
    return torch.empty([int(s) for s in shape], dtype=dtype, device=_DEVICE if _DEVICE is not None else device())


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t.numel() else ctypes.c_void_p(0)


def synchronize():
    torch.cuda.synchronize(device())


_SIDE = {}


def side_stream():
    """A second stream (default priority) on the current device, for setup work that runs beside the
    caller's stream: fork() starts it after everything the caller's stream has queued, join() makes the
    caller's stream wait for it."""
    d = device()
    s = _SIDE.get(d.index)
    if s is None:
        s = _SIDE[d.index] = torch.cuda.Stream(d)
    return s


def fork():
    side = side_stream()
    side.wait_stream(torch.cuda.current_stream())
    return side


def join(side, produced=()):
    """The caller's stream waits for `side`; the device tensors of the objects in `produced` (made on
    `side`, used on the caller's stream from here on) are recorded as in use on the caller's stream, so
    that the allocator does not hand their memory to new work on `side` before that use is done."""
    main = torch.cuda.current_stream()
    main.wait_stream(side)
    for obj in produced:
        for v in (vars(obj).values() if hasattr(obj, "__dict__") else ()):
            if isinstance(v, torch.Tensor) and v.is_cuda:
                v.record_stream(main)
