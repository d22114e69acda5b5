"""Device-buffer helpers for the GPU tests: numpy limb arrays <-> torch tensors."""
import numpy as np
import torch

_VIEW = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}
_BACK = {torch.int32: np.uint32, torch.int64: np.uint64}


def to_dev(a):
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(_VIEW[a.dtype]).copy()).cuda()


def empty(n, bits):
    return torch.zeros(max(int(n), 1), dtype=torch.int64 if bits == 64 else torch.int32, device="cuda")


def from_dev(t, n=None):
    torch.cuda.synchronize()
    a = t.cpu().numpy().view(_BACK[t.dtype])
    return a if n is None else a[:n]
