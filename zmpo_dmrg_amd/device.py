"""Device memory and stream plumbing (PyTorch-ROCm allocations, raw pointers).

All numerics run in libzmpo kernels; torch only owns the allocations and the
current HIP stream.
"""
import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def available():
    return torch().cuda.is_available()


def require():
    if not available():
        from ._lib import ZmpoError
        raise ZmpoError('no HIP device visible: the zmpo_dmrg_amd device path needs a GPU')


def device():
    return torch().device('cuda')


TORCH_DTYPES = {0: 'float64', 1: 'float32'}


def tdtype(code):
    return getattr(torch(), TORCH_DTYPES[code])


def empty(n, code=0):
    require()
    return torch().empty(int(max(n, 1)), dtype=tdtype(code), device=device())


def zeros(n, code=0):
    require()
    return torch().zeros(int(max(n, 1)), dtype=tdtype(code), device=device())


def upload(arr, code=None):
    """Host numpy -> device tensor (float64/float32/int64 by dtype)."""
    require()
    t = torch()
    a = np.ascontiguousarray(arr)
    if code is not None:
        a = a.astype(np.float64 if code == 0 else np.float32)
    return t.from_numpy(a).to(device())


def upload_bytes(struct_arr):
    """Structured numpy array -> device uint8 tensor holding its bytes."""
    require()
    t = torch()
    raw = np.ascontiguousarray(struct_arr).view(np.uint8).reshape(-1)
    if raw.size == 0:
        raw = np.zeros(8, dtype=np.uint8)
    return t.from_numpy(raw.copy()).to(device())


def ptr(x):
    return None if x is None else x.data_ptr()


def stream():
    return torch().cuda.current_stream().cuda_stream


def stream_obj():
    """The torch stream object whose handle stream() returns (kernel launch stream)."""
    return torch().cuda.current_stream()


def sync():
    torch().cuda.synchronize()


def download(x):
    return x.detach().cpu().numpy()
