"""Device buffers and streams for the C ABI.

PyTorch-ROCm is used only as plumbing here: it owns HBM allocations (its
caching allocator), host<->device copies, and the current HIP stream whose
handle is passed to every ``frmc_*`` call.  No computation of the path runs
in PyTorch.
"""
import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t
        _torch = _t
    return _torch


def require_gpu():
    """Raise unless a ROCm GPU is visible: there is no CPU code path."""
    t = torch()
    if not t.cuda.is_available():
        from ._lib import FrmcError
        raise FrmcError("fullrmc_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                        "the HIP kernels have no CPU fallback")
    return t


def stream():
    """Handle of torch's current HIP stream, as the void* the C ABI takes."""
    t = require_gpu()
    return t.cuda.current_stream().cuda_stream


_NP2T = {np.dtype(np.float32): "float32", np.dtype(np.int32): "int32", np.dtype(np.float64): "float64",
         np.dtype(np.uint8): "uint8"}


def upload(a):
    """Copy a host numpy array to a new device tensor (C-contiguous)."""
    t = require_gpu()
    a = np.ascontiguousarray(a)
    host = t.from_numpy(a) if a.size else t.empty(a.shape, dtype=getattr(t, _NP2T[a.dtype]))
    return host.to("cuda", non_blocking=False)


def empty(shape, dtype):
    t = require_gpu()
    return t.empty(shape, dtype=getattr(t, _NP2T[np.dtype(dtype)]), device="cuda")


def zeros(shape, dtype):
    t = require_gpu()
    return t.zeros(shape, dtype=getattr(t, _NP2T[np.dtype(dtype)]), device="cuda")


def download(tensor):
    """Device tensor -> new host numpy array (synchronises the stream)."""
    return tensor.cpu().numpy()


def ptr(tensor):
    return tensor.data_ptr() if tensor is not None else None
