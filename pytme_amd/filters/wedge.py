"""Reconstructed missing-wedge masks (``tme/filters/wedge.py:271-507``).

The step wedge is the support of the Fourier planes sampled by each tilt:
a central line of the (opening, tilt) plane is rotated to every tilt angle
with order-1 interpolation, the lines are summed, clipped to the maximal
weight and tiled along the remaining axis.  The continuous wedge is the
analytic double cone between the extreme tilt angles.  Both are binarised
(``weight_wedge=False``), limited to the frequency cutoff, moved to the FFT
origin and cropped to the rfft half spectrum.
"""

from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np
from scipy.ndimage import affine_transform

from ..matching_utils import centered
from ..rotations import euler_to_rotationmatrix
from ._utils import centered_grid, crop_real_fourier, fftfreqn, shift_fourier

__all__ = ["WedgeReconstructed"]


def _rotate_plane_order1(plane: np.ndarray, rotation_2d: np.ndarray) -> np.ndarray:
    """Rotate a 2D plane about its geometric centre (scipy order 1, constant).

    The float32 homogeneous pull matrix T(c) R^-1 T(-c) is assembled with
    float32 products as the reference numpy backend does
    (npfftw_backend.py:326-388).
    """
    ndim = 2
    center = np.divide(np.asarray(plane.shape) - 1, 2).astype(np.float32)
    matrix = np.eye(ndim + 1, dtype=np.float32)
    tmat = np.eye(ndim + 1, dtype=np.float32)
    tmat[:ndim, ndim] = 0
    matrix = np.dot(matrix, tmat)
    cmat = np.eye(ndim + 1, dtype=np.float32)
    cmat[:ndim, ndim] = center
    matrix = np.dot(matrix, cmat)
    rmat = np.eye(ndim + 1, dtype=np.float32)
    rmat[:ndim, :ndim] = np.linalg.inv(rotation_2d.astype(np.float32))
    matrix = np.dot(matrix, rmat)
    cmat[:ndim, ndim] = -cmat[:ndim, ndim]
    matrix = np.dot(matrix, cmat)
    matrix /= matrix[ndim, ndim]
    out = np.zeros_like(plane)
    affine_transform(plane, matrix, mode="constant", output=out, order=1, prefilter=True)
    return out


@dataclass
class WedgeReconstructed:
    """Wedge mask of a tilt series reconstruction.

    Parameters mirror the reference dataclass: ``angles`` (degrees),
    ``weights``, ``weight_wedge``, ``create_continuous_wedge``,
    ``frequency_cutoff`` (cycles/voxel), ``tilt_axis`` and ``opening_axis``.
    """

    angles: Tuple[float] = None
    weights: Tuple[float] = None
    weight_wedge: bool = False
    create_continuous_wedge: bool = False
    frequency_cutoff: float = 0.5
    tilt_axis: int = 0
    opening_axis: int = 2
    reconstruction_filter: str = None

    def __post_init__(self):
        if self.create_continuous_wedge:
            self.angles = (min(self.angles), max(self.angles))
        if self.reconstruction_filter is not None:
            raise NotImplementedError("reconstruction filters are not part of this build")

    def __call__(self, shape: Tuple[int], return_real_fourier: bool = False, **kwargs) -> Dict:
        args = dict(vars(self))
        args.update(kwargs)
        func = self.continuous_wedge if args.get("create_continuous_wedge", False) else self.step_wedge
        weight_wedge = args.get("weight_wedge", False)
        if args.get("wedge_weights") is None and weight_wedge:
            args["weights"] = np.cos(np.radians(np.asarray(args.get("angles", (0,)))))
        ret = func(shape=shape, **{k: v for k, v in args.items() if k != "shape"})
        cutoff = args.get("frequency_cutoff", None)
        if cutoff is not None:
            freq = fftfreqn(shape=shape, sampling_rate=1, compute_euclidean_norm=True)
            ret = np.multiply(ret, freq <= cutoff, out=ret)
        if not weight_wedge:
            ret = (ret > 0) * 1.0
        ret = shift_fourier(np.asarray(ret, dtype=np.float32))
        if return_real_fourier:
            ret = crop_real_fourier(ret)
        return {
            "data": np.ascontiguousarray(ret),
            "shape": shape,
            "return_real_fourier": return_real_fourier,
            "is_multiplicative_filter": True,
        }

    @staticmethod
    def continuous_wedge(shape, angles, opening_axis: int, tilt_axis: int, **kwargs) -> np.ndarray:
        """Double cone |k_tilt / k_opening| beyond the extreme tilt angles."""
        angles = np.abs(np.asarray(angles, dtype=np.float64))
        aspect = shape[opening_axis] / shape[tilt_axis]
        angles = np.degrees(np.arctan(np.tan(np.radians(angles)) * aspect))
        start = np.tan(np.radians(90 - angles[0]))
        stop = np.tan(np.radians(-1 * (90 - angles[1])))
        grid = centered_grid(shape)
        with np.errstate(divide="ignore", invalid="ignore"):
            ratios = np.where(
                grid[opening_axis] == 0, np.tan(np.radians(90)) + 1, grid[tilt_axis] / grid[opening_axis]
            )
        return np.logical_or(start <= ratios, stop >= ratios).astype(np.float32)

    @staticmethod
    def step_wedge(shape, angles, opening_axis: int, tilt_axis: int, weights=None, **kwargs) -> np.ndarray:
        """Sum of per-tilt central lines of the (opening, tilt) plane."""
        angles = np.asarray(angles, dtype=np.float64)
        weights = np.ones(angles.size) if weights is None else np.asarray(weights)
        weights = np.repeat(weights, angles.size // weights.size)
        shape = tuple(int(x) for x in shape)
        opening_axis, tilt_axis = int(opening_axis), int(tilt_axis)
        plane = np.zeros((shape[opening_axis], shape[tilt_axis] + (1 - shape[tilt_axis] % 2)), dtype=np.float32)
        aspect = plane.shape[0] / plane.shape[1]
        angles = np.degrees(np.arctan(np.tan(np.radians(angles)) * aspect))
        row = plane.shape[0] // 2
        wedge = np.zeros_like(plane)
        for i in range(angles.shape[0]):
            plane[row : row + 1, :] = 1
            rot = euler_to_rotationmatrix((angles[i], 0))[np.ix_((0, 1), (0, 1))]
            wedge += _rotate_plane_order1(plane * 1, rot) * weights[i]
        wedge = centered(wedge, (shape[opening_axis], shape[tilt_axis]))
        np.fmin(wedge, np.max(weights), wedge)
        if opening_axis > tilt_axis:
            wedge = np.moveaxis(wedge, 1, 0)
        dims = tuple(x if i in (opening_axis, tilt_axis) else 1 for i, x in enumerate(shape))
        wedge = wedge.reshape(dims)
        return np.tile(wedge, np.divide(shape, dims).astype(int))
