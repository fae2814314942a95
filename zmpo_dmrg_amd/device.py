"""Device memory and stream plumbing (PyTorch-ROCm allocations, raw pointers).

All numerics run in libzmpo kernels; torch only owns the allocations and the
current HIP stream.
"""
import os

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        # environments and intermediates of many sizes (up to tens of GB)
        # are allocated and freed every step: cached blocks above 256 MB are
        # never split, so a large request either reuses a whole block or
        # makes torch return its free cached blocks to the device before
        # allocating (growable segments are not available on this
        # platform).  Read when the allocator starts: it applies unless the
        # process set it or allocated device memory first.
        os.environ.setdefault('PYTORCH_ALLOC_CONF', 'max_split_size_mb:256')
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


# storage of the kernel dtype codes: 0 FP64, 1 FP32, 2 BF16 (ZMPO_BF16),
# 3 FP32 (the output / workspace side of ZMPO_BF16_F32)
TORCH_DTYPES = {0: 'float64', 1: 'float32', 2: 'bfloat16', 3: 'float32'}


def tdtype(code):
    return getattr(torch(), TORCH_DTYPES[code])


def empty(n, code=0):
    require()
    return torch().empty(int(max(n, 1)), dtype=tdtype(code), device=device())


def zeros(n, code=0):
    """n zeros of storage type code on the launch stream."""
    return clear(empty(n, code), code)


def clear(x, code=None):
    """x[:] = 0: FP64 / FP32 by the libzmpo lincomb kernel (no operands,
    beta = 0) on the launch stream; other storage types by torch."""
    if code is None:
        code = {torch().float64: 0, torch().float32: 1}.get(x.dtype, -1)
    if code in (0, 1):
        from . import _lib
        _lib.check(_lib.lib().zmpo_lincomb(code, ptr(x), None, 0, 0, None, 1.0, 0.0, x.numel(),
                                           stream()))
    else:
        x.zero_()
    return x


ELSIZE = {0: 8, 1: 4, 2: 2, 3: 4}
_SCRATCH = {}


def scratch(name, nbytes):
    """A device byte buffer of at least nbytes, one per name, reused by
    every caller (grown when too small, never shrunk): the sigma and
    renormalisation intermediates of all operand types share one, so the
    precision tiers and successive steps do not reallocate them."""
    require()
    key = (str(device()), name)
    buf = _SCRATCH.get(key)
    if buf is None or buf.numel() < nbytes:
        _SCRATCH.pop(key, None)        # release the old buffer first
        buf = None
        buf = torch().empty(int(max(nbytes, 256)), dtype=torch().uint8, device=device())
        _SCRATCH[key] = buf
    return buf


def typed(buf, off_bytes, n, code):
    """n elements of dtype code viewed at byte offset off_bytes (a multiple
    of 8) of a scratch buffer."""
    es = ELSIZE[code]
    return buf[off_bytes:off_bytes + int(max(n, 1)) * es].view(tdtype(code))


def release_scratch():
    """Drop every scratch buffer (the memory returns to the allocator)."""
    _SCRATCH.clear()


def upload(arr, code=None):
    """Host numpy -> device tensor (float64/float32/int64 by dtype)."""
    require()
    t = torch()
    a = np.ascontiguousarray(arr)
    if code is not None:
        a = a.astype(np.float64 if code == 0 else np.float32)
    out = t.from_numpy(a).to(device())
    # BF16: rounded to nearest even on the device
    return out.to(t.bfloat16) if code == 2 else out


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
    x = x.detach()
    if x.dtype == torch().bfloat16:
        x = x.float()
    return x.cpu().numpy()
