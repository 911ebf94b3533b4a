"""Rotation conventions of the scan (subset of ``tme/rotations.py``).

Rotation matrices are float32 (n, 3, 3) "push" rotations; the scan applies
their inverse as the pull map (tme/matching_data.py:641-660,
tme/backends/npfftw_backend.py:326-355).
"""

from typing import Tuple

import numpy as np
from scipy.spatial.transform import Rotation

__all__ = [
    "euler_to_rotationmatrix",
    "euler_from_rotationmatrix",
    "uniform_rotations",
    "get_rotation_matrices",
]


def euler_to_rotationmatrix(angles: Tuple[float], seq: str = "zyz") -> np.ndarray:
    """Rotation matrix of intrinsic/extrinsic Euler ``angles`` (degrees).

    Uses scipy's ``Rotation.from_euler(seq[:len(angles)], degrees=True)`` and
    returns float32, as tme/rotations.py:171-196.
    """
    angles = np.asarray(angles)
    n = angles.shape[1] if angles.ndim == 2 else len(angles)
    return Rotation.from_euler(seq=seq[:n], angles=angles, degrees=True).as_matrix().astype(np.float32)


def euler_from_rotationmatrix(rotation_matrix: np.ndarray, seq: str = "zyz") -> np.ndarray:
    """Euler angles (degrees, float32) of a 2x2 or 3x3 rotation."""
    rotation_matrix = np.asarray(rotation_matrix)
    if rotation_matrix.shape[0] == 2:
        tmp = np.eye(3)
        tmp[:2, :2] = rotation_matrix
        rotation_matrix = tmp
    return Rotation.from_matrix(rotation_matrix).as_euler(seq=seq, degrees=True).astype(np.float32)


def uniform_rotations(n: int, seed: int = 0, include_identity: bool = False) -> np.ndarray:
    """``n`` rotations uniform on SO(3) from a seeded generator.

    Quaternions q ~ N(0, I4) normalised, sign chosen so the scalar part is
    non-negative, interpreted scalar-last as in
    ``Rotation.from_quat`` (tme/rotations.py:253); returned as float32.
    """
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((int(n), 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 3] < 0] *= -1
    mats = Rotation.from_quat(q).as_matrix().astype(np.float32)
    if include_identity and n > 0:
        mats[0] = np.eye(3, dtype=np.float32)
    return mats


def get_rotation_matrices(angular_sampling: float, dim: int = 3, seed: int = 0) -> np.ndarray:
    """Random rotation set whose size follows the requested sampling.

    The reference's optimised quaternion coverings (tme/data/*.npy) are not
    part of this package; a uniform random set of comparable size is
    returned instead, identity first.
    """
    k = max(int((360.0 / float(angular_sampling)) ** (dim * (dim - 1) // 2) / 8), 1)
    if dim == 3:
        return uniform_rotations(k, seed=seed, include_identity=True)
    rng = np.random.default_rng(seed)
    mats, _ = np.linalg.qr(rng.standard_normal((k, dim, dim)))
    dets = np.linalg.det(mats)
    mats[dets < 0, :, -1] *= -1
    mats[0] = np.eye(dim)
    return mats.astype(np.float32)
