"""Shared helpers: column-major numpy <-> device tensors, seeded inputs."""
import numpy as np


def dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.copy())
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def empty(n, dtype=None):
    import torch
    return torch.empty(int(n), dtype=dtype or torch.float32, device="cuda")


def fcol(M):
    """2-D array -> flat column-major float32"""
    return np.ascontiguousarray(np.asarray(M, np.float32).ravel(order="F"))


def mat(flat, rows, cols):
    return np.asarray(flat).reshape((rows, cols), order="F")


def relnorm(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
