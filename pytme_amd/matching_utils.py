"""Utilities of the exhaustive scan (subset of ``tme/matching_utils.py``)."""

from itertools import product
from typing import Callable, Dict, Optional, Tuple

import numpy as np

from .backends import backend as be

__all__ = [
    "noop",
    "identity",
    "conditional_execute",
    "normalize_template",
    "centered",
    "apply_convolution_mode",
    "split_shape",
    "elliptical_mask",
]


def noop(*args, **kwargs):
    return None


def identity(arr, *args):
    return arr


def conditional_execute(func: Callable, execute_operation: bool, alt_func: Callable = noop) -> Callable:
    """``func`` if ``execute_operation`` else ``alt_func`` (tme/matching_utils.py:36-58)."""
    return func if execute_operation else alt_func


def normalize_template(template, mask, n_observations, axis=None):
    """Standardise ``template`` to zero mean, unit std inside ``mask`` (in place).

    mean = sum(t*m)/n,  std = sqrt(max(sum(t^2*m)/n - mean^2, 0)),
    t <- (t - mean)/std * m   (tme/matching_utils.py:61-99).

    Backends that provide ``normalize_template`` (the hip backend: FP64
    moments in one deterministic reduction) run it natively.
    """
    native = getattr(be, "normalize_template", None)
    if native is not None:
        ret = native(template, mask, n_observations, axis=axis)
        if ret is not None:
            return ret
    masked_mean = be.sum(be.multiply(template, mask), axis=axis, keepdims=True)
    masked_mean = be.divide(masked_mean, n_observations)
    masked_std = be.sum(be.multiply(be.square(template), mask), axis=axis, keepdims=True)
    masked_std = be.subtract(be.divide(masked_std, n_observations), be.square(masked_mean))
    masked_std = be.sqrt(be.maximum(masked_std, 0))
    template = be.subtract(template, masked_mean, out=template)
    template = be.divide(template, masked_std, out=template)
    return be.multiply(template, mask, out=template)


def _normalize_template_overflow_safe(template, mask, n_observations, axis=None):
    """Variant for 2-byte dtypes: normalise in the overflow-safe dtype."""
    _template = be.astype(template, be._overflow_safe_dtype)
    _mask = be.astype(mask, be._overflow_safe_dtype)
    normalize_template(_template, _mask, n_observations, axis=axis)
    template[:] = be.astype(_template, template.dtype)
    return template


def _center_slice(current_shape: Tuple[int], new_shape: Tuple[int]) -> Tuple[slice]:
    starts = [(int(c) - int(n)) // 2 for c, n in zip(current_shape, new_shape)]
    return tuple(slice(s, s + int(n)) for s, n in zip(starts, new_shape))


def centered(arr, new_shape: Tuple[int]):
    """Central ``new_shape`` box of ``arr`` (scipy.signal's ``_centered``)."""
    return arr[_center_slice(arr.shape, new_shape)]


def centered_mask(arr, new_shape: Tuple[int]):
    """Zero everything outside the central ``new_shape`` box."""
    box = _center_slice(arr.shape, new_shape)
    mask = np.zeros(arr.shape, dtype=bool)
    mask[box] = True
    arr[~mask] = 0
    return arr


def apply_convolution_mode(
    arr,
    convolution_mode: str,
    s1: Tuple[int],
    s2: Tuple[int],
    convolution_shape: Tuple[int] = None,
    mask_output: bool = False,
):
    """Crop a convolution result to ``full``/``same``/``valid``.

    First drops the padding to the fast FFT length (crop to
    ``convolution_shape``), then takes the centred ``s1`` (same) or
    ``s1 - s2 + 1`` (valid) box (tme/matching_utils.py:469-524).
    """
    if convolution_shape is None:
        convolution_shape = [int(a) + int(b) - 1 for a, b in zip(s1, s2)]
    arr = arr[tuple(slice(0, int(x)) for x in convolution_shape)]
    if convolution_mode not in ("full", "same", "valid"):
        raise ValueError("Supported convolution_mode are 'full', 'same' and 'valid'.")
    func = centered_mask if mask_output else centered
    if convolution_mode == "full":
        return arr
    if convolution_mode == "same":
        return func(arr, s1)
    return func(arr, [int(a) - int(b) + 1 for a, b in zip(s1, s2)])


def split_shape(shape: Tuple[int], splits: Dict, equal_shape: bool = True) -> Tuple:
    """Split ``shape`` into (possibly overlapping) equally sized boxes.

    ``splits`` maps axis -> number of pieces.  All pieces but the last along
    an axis start at multiples of ceil(n / k); with ``equal_shape`` the last
    piece is aligned to the end so every piece has the same length
    (tme/matching_utils.py:573-618).
    """
    ndim = len(shape)
    splits = {k: max(int(splits.get(k, 1)), 1) for k in range(ndim)}
    lengths = np.divide(shape, [splits[i] for i in range(ndim)])
    lengths = np.ceil(lengths).astype(int) if equal_shape else lengths.astype(int)
    per_axis = []
    for axis in range(ndim):
        n, length, k = int(shape[axis]), int(lengths[axis]), splits[axis]
        pieces = []
        for i in range(k):
            if i < k - 1:
                pieces.append(slice(i * length, (i + 1) * length))
            elif equal_shape:
                pieces.append(slice(n - length, n))
            else:
                pieces.append(slice(i * length, n))
        per_axis.append(tuple(pieces))
    return tuple(product(*per_axis))


def elliptical_mask(
    shape: Tuple[int],
    radius: Tuple[float],
    center: Optional[Tuple[float]] = None,
    **kwargs,
) -> np.ndarray:
    """Binary ellipsoid: ``||(x - center) / radius|| <= 1``.

    ``center`` defaults to ``shape // 2`` (tme/matching_utils.py:765-841,
    without orientation or soft edge).
    """
    shape = np.asarray(shape, dtype=int)
    radius = np.asarray(radius, dtype=np.float64)
    if center is None:
        center = np.divide(shape, 2).astype(int)
    center = np.asarray(center, dtype=np.float32)
    radius = np.repeat(radius, shape.size // radius.size)
    center = np.repeat(center, shape.size // center.size)
    if radius.size != shape.size:
        raise ValueError("Length of radius has to be either one or match shape.")
    if center.size != shape.size:
        raise ValueError("Length of center has to be either one or match shape.")
    grid = np.indices(tuple(shape), dtype=np.float32)
    grid = grid - center.reshape((-1,) + (1,) * shape.size)
    dist = np.linalg.norm(grid / radius.reshape((-1,) + (1,) * shape.size), axis=0)
    return (dist <= 1).astype(int)
