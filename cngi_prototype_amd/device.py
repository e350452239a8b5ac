"""Device-memory plumbing: move arrays to the GPU and back.

Arrays held by a dataset may be numpy arrays (host) or torch tensors already
resident in HBM; ``to_device`` returns a contiguous device tensor either way
without copying device-resident data.
"""
import numpy as np


def to_device(x, dtype=None):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible; the MI355X gridding path needs a GPU")
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.cuda()
    else:
        a = np.ascontiguousarray(np.asarray(x))
        t = torch.from_numpy(a).cuda()
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def to_host(t):
    import torch
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return np.asarray(t)
