"""Frequency grids and spectrum layout helpers (``tme/filters/_utils.py``).

All grids are float32 and computed on the host, matching the reference's
float32 grid arithmetic.
"""

from typing import Tuple

import numpy as np

__all__ = ["fftfreqn", "crop_real_fourier", "shift_fourier", "centered_grid", "compute_tilt_shape"]


def compute_tilt_shape(shape: Tuple[int], opening_axis: int, reduce_dim: bool = False):
    if reduce_dim:
        return tuple(x for i, x in enumerate(shape) if i != opening_axis)
    return tuple(x if i != opening_axis else 1 for i, x in enumerate(shape))


def centered_grid(shape: Tuple[int]) -> np.ndarray:
    """Integer grid with the origin at ``size // 2`` along every axis."""
    return np.array(np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij"))


def fftfreqn(
    shape: Tuple[int],
    sampling_rate,
    compute_euclidean_norm: bool = False,
    shape_is_real_fourier: bool = False,
    return_sparse_grid: bool = False,
):
    """Centred sample frequencies ``(i - n//2) / (n * sampling_rate)`` (float32).

    With ``sampling_rate=None`` the grid is the plain index offset.  With
    ``shape_is_real_fourier`` the last axis runs 0..n-1 over the rfft half
    spectrum of length 2*(n-1).
    """
    shape = tuple(int(x) for x in shape)
    ndim = len(shape)
    center = np.floor_divide(shape, 2).astype(np.int32)
    norm = np.ones(ndim, dtype=np.float32)
    if sampling_rate is not None:
        norm = np.multiply(shape, sampling_rate).astype(int)
    if shape_is_real_fourier:
        center[-1] = 0
        norm = norm.astype(np.float64)
        norm[-1] = 1 if sampling_rate is None else (shape[-1] - 1) * 2 * np.max(sampling_rate)
    grids = []
    for i, n in enumerate(shape):
        dims = tuple(1 if i != t else n for t in range(ndim))
        g = ((np.arange(n, dtype=np.int32) - center[i]) / norm[i]).astype(np.float32)
        grids.append(g.reshape(dims))
    if compute_euclidean_norm:
        out = sum(np.square(g) for g in grids)
        return np.sqrt(out, out=out)
    if return_sparse_grid:
        return grids
    flesh = np.ones(shape, dtype=np.float32)
    return np.stack([g * flesh for g in grids])


def crop_real_fourier(data):
    """Keep the rfft half spectrum of the last axis."""
    return data[..., : 1 + data.shape[-1] // 2]


def shift_fourier(data, shape_is_real_fourier: bool = False):
    """Move the DC term from the array centre to index 0 (ifftshift)."""
    shift = [int(n // 2 + n % 2) for n in data.shape]
    if shape_is_real_fourier:
        shift[-1] = 0
    return np.roll(data, shift, axis=tuple(range(len(shift))))
