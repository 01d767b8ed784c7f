"""Leaf-stage chunk sizes (``_device_config.py:37-90`` of the reference).

The build here never chunks the leaf stage: every leaf's factors stay
resident in the GPU's HBM (288 GB on MI355X).  These functions keep the
reference's API and report how many leaves of degree p fit, rounded down to
a power of four (2D) or eight (3D), from the device's free memory (the full
MI355X capacity when no GPU is visible).
"""

from __future__ import annotations

import torch

_HBM_BYTES = 288 * 1024**3


def _free_bytes() -> int:
    if torch.cuda.is_available():
        free, _ = torch.cuda.mem_get_info()
        return int(free)
    return _HBM_BYTES


def _elem(dtype) -> int:
    try:
        return torch.empty((), dtype=dtype).element_size()
    except TypeError:
        import numpy as np

        return np.dtype(dtype).itemsize


def _leaves_that_fit(n: int, nq: int, dtype, base: int) -> int:
    # leaf factors n^2, Y n x nq, T nq^2, workspace ~ 2 n (nq + 1): per leaf
    per_leaf = (2 * n * n + n * nq + nq * nq + 2 * n * (nq + 1)) * _elem(dtype)
    count = max(1, int(0.5 * _free_bytes() // per_leaf))
    out = 1
    while out * base <= count:
        out *= base
    return out


def local_solve_chunksize_2D(p: int, dtype) -> int:
    """Leaves of a 2D degree-p problem whose leaf stage fits at once (a power of 4)."""
    return _leaves_that_fit(p * p, 4 * (p - 2), dtype, 4)


def local_solve_chunksize_3D(p: int, dtype) -> int:
    """Leaves of a 3D degree-p problem whose leaf stage fits at once (a power of 8)."""
    return _leaves_that_fit(p**3, 6 * (p - 2) ** 2, dtype, 8)
