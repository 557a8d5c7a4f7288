"""Helpers for the GPU parity tests: numpy <-> device tensors, relative errors."""
import sys

import numpy as np


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(*args):
    """Names the case on stdout before a GPU step (flushed)."""
    print(*args, flush=True)
    sys.stdout.flush()
