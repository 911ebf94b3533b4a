"""Coordinate maps between target and padded score space.

Host-side (numpy) restatement of ``tme/analyzer/_utils.py:17-191``.
"""

from typing import Tuple

import numpy as np

__all__ = ["cart_to_score", "score_to_cart"]


def _convmode_to_shape(convolution_mode, targetshape, templateshape, convolution_shape):
    if convolution_mode == "same":
        return np.asarray(targetshape, dtype=int)
    if convolution_mode == "valid":
        return np.subtract(targetshape, templateshape) + 1
    return np.asarray(convolution_shape, dtype=int)


def cart_to_score(
    positions,
    fast_shape: Tuple[int],
    targetshape: Tuple[int],
    templateshape: Tuple[int],
    convolution_shape: Tuple[int] = None,
    fourier_shift: Tuple[int] = None,
    convolution_mode: str = None,
    **kwargs,
):
    """Target coordinates -> raw (circular, fast_shape) score coordinates."""
    positions = np.asarray(positions, dtype=int)
    out_shape = _convmode_to_shape(convolution_mode, targetshape, templateshape, convolution_shape)
    valid = np.all((positions >= 0) & (positions < out_shape), axis=1)
    starts = np.trunc(np.subtract(convolution_shape, out_shape) / 2).astype(int)
    positions = positions + starts
    if fourier_shift is not None:
        positions = np.mod(positions - np.asarray(fourier_shift, dtype=int), np.asarray(fast_shape, dtype=int))
    return positions, valid


def score_to_cart(
    positions,
    fast_shape: Tuple[int] = None,
    targetshape: Tuple[int] = None,
    templateshape: Tuple[int] = None,
    convolution_shape: Tuple[int] = None,
    fourier_shift: Tuple[int] = None,
    convolution_mode: str = None,
    **kwargs,
):
    """Raw score coordinates -> target coordinates, with a validity flag."""
    positions = np.asarray(positions, dtype=int)
    valid = np.ones(positions.shape[0], dtype=bool)
    if fourier_shift is not None:
        positions = np.mod(positions + np.asarray(fourier_shift, dtype=int), np.asarray(fast_shape, dtype=int))
    if convolution_mode is not None:
        out_shape = _convmode_to_shape(convolution_mode, targetshape, templateshape, convolution_shape)
        starts = np.trunc(np.subtract(convolution_shape, out_shape) / 2).astype(int)
        stops = starts + out_shape
        valid = np.all((positions >= starts) & (positions < stops), axis=1)
        positions = positions - starts
    valid &= np.all(positions >= 0, axis=1)
    return positions, valid
