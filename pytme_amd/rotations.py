"""Rotation conventions of the scan (subset of ``tme/rotations.py``).

Rotation matrices are float32 (n, 3, 3) "push" rotations; the scan applies
their inverse as the pull map (tme/matching_data.py:641-660,
tme/backends/npfftw_backend.py:326-355).
"""

import importlib.util
import os
from typing import Optional, Tuple

import numpy as np
from scipy.spatial.transform import Rotation

__all__ = [
    "euler_to_rotationmatrix",
    "euler_from_rotationmatrix",
    "uniform_rotations",
    "get_rotation_matrices",
    "load_quaternions_by_angle",
    "rotation_set_directory",
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


def rotation_set_directory(data_dir: Optional[str] = None) -> str:
    """Directory holding pyTME's orientation sets (``metadata.yaml`` and the
    ``c48*.npy`` quaternion files of ``tme/data``): ``data_dir``, else the
    environment variable ``PYTME_ROTATION_SETS``, else the ``data`` folder
    of an installed ``tme`` package (located without importing it).  Raises
    ``FileNotFoundError`` when none holds ``metadata.yaml``."""
    candidates = [data_dir, os.environ.get("PYTME_ROTATION_SETS")]
    spec = importlib.util.find_spec("tme") if data_dir is None else None
    if spec is not None and spec.submodule_search_locations:
        candidates += [os.path.join(p, "data") for p in spec.submodule_search_locations]
    for path in candidates:
        if path and os.path.isfile(os.path.join(path, "metadata.yaml")):
            return path
    raise FileNotFoundError(
        "no pyTME orientation sets found: pass data_dir (a copy of tme/data with metadata.yaml and the "
        "c48*.npy files), set PYTME_ROTATION_SETS, or install pyTME; uniform_rotations() gives a seeded "
        "random set instead"
    )


def load_quaternions_by_angle(angular_sampling: float, data_dir: Optional[str] = None):
    """The orientation set whose sampling is closest to ``angular_sampling``
    (tme/rotations.py:266-300): ``(quaternions (n, 4), weights (n,),
    set sampling in degrees)``.  ``metadata.yaml`` maps each file to
    (orientations, sampling, coverage); files are read with
    ``numpy.load`` (no pickles)."""
    import yaml

    path = rotation_set_directory(data_dir)
    with open(os.path.join(path, "metadata.yaml")) as fh:
        meta = yaml.safe_load(fh)
    name = min(meta, key=lambda k: abs(float(angular_sampling) - float(meta[k][1])))
    table = np.load(os.path.join(path, name))
    return table[:, :4], table[:, -1], meta[name][0]


def get_rotation_matrices(angular_sampling: float, dim: int = 3, use_optimized_set: bool = True,
                          data_dir: Optional[str] = None, seed: Optional[int] = None) -> np.ndarray:
    """Rotation matrices at ``angular_sampling`` degrees (tme/rotations.py:227-263).

    ``dim == 3`` with ``use_optimized_set``: the closest pyTME orientation
    set (:func:`load_quaternions_by_angle`), its quaternions read
    scalar-last as the reference passes them to ``Rotation.from_quat``
    (the reference's quirk is kept so that both scan the same matrices);
    raises ``FileNotFoundError`` when the sets are not available.
    Otherwise ``(360 / angular_sampling) ** (d (d - 1) / 2)`` random
    orthogonal matrices with determinant 1 from QR decompositions, identity
    first (``seed`` fixes the generator; the reference draws from numpy's
    global one).
    """
    if dim == 3 and use_optimized_set:
        quaternions, _, _ = load_quaternions_by_angle(angular_sampling, data_dir)
        return Rotation.from_quat(quaternions).as_matrix()
    k = int((360 / angular_sampling) ** (dim * (dim - 1) // 2))
    rng = np.random.default_rng(seed) if seed is not None else np.random
    draws = rng.standard_normal((k, dim, dim)) if seed is not None else rng.randn(k, dim, dim)
    mats, _ = np.linalg.qr(draws)
    mats[np.linalg.det(mats) < 0, :, -1] *= -1
    mats[0] = np.eye(dim)
    return mats
