"""Helpers shared by the GPU tests: column-major matrices as torch device tensors."""
import numpy as np


def to_dev(A, dtype=None):
    """Column-major numpy (m x n) -> torch CUDA tensor of shape (n, m) (same memory order, ld = m)."""
    import torch
    a = np.ascontiguousarray(np.asarray(A).T)
    t = torch.from_numpy(a)
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def from_dev(T):
    """Inverse of to_dev: torch (n, m) -> numpy column-major (m x n) FP64."""
    return np.asfortranarray(T.detach().to("cpu").double().numpy().T)


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
