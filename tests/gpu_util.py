"""Helpers for the GPU parity tests: numpy <-> device via torch (plumbing)."""
import numpy as np


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype.fields is not None:
        a = a.view(np.int32)
    return torch.from_numpy(a.copy()).to("cuda")


def host(t, dtype=np.float32, shape=None):
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy().view(dtype)
    return a.reshape(shape) if shape is not None else a


def ptr(t):
    return None if t is None else t.data_ptr()


def stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def empty(n, fill=np.nan):
    import torch
    return torch.full((int(n),), float(fill), dtype=torch.float32, device="cuda")
