"""Peak calling and local refinement of a scan result without leaving the
GPU (the peak-calling and ``--local-optimization`` steps of
scripts/postprocess.py:482-613).

The reference moves the max-score volume to the host, calls peaks there,
looks every peak's rotation up in the rotation-index volume and the
rotation mapping, converts it to zyz Euler angles and then runs a
derivative-free optimiser per peak.  Here the analyzer's state stays in
HBM: it is rolled and cropped on the device exactly as
``MaxScoreOverRotations.result`` does, the GPU peak caller folds it, the
peaks' rotation indices are gathered from the rotation-index volume and
their matrices from the scan's rotation table, the matrices become zyz
angles in FP64 on the device, and all peaks are refined at once by the
gradient refinement (:func:`~pytme_amd.refine.refine_peaks`).  Only the
final poses come back to the host.
"""

from typing import Dict, Optional, Tuple

import numpy as np

__all__ = ["device_result", "rotation_table", "rotation_matrix_to_euler_zyz", "peak_poses", "call_and_refine"]


def _torch():
    import torch

    return torch


def device_result(state, fourier_shift=None, convolution_mode=None, targetshape=None, templateshape=None,
                  convolution_shape=None, **kwargs):
    """The analyzer state ``(scores, rotation_indices, mapping)`` rolled by
    ``fourier_shift`` and cropped by ``convolution_mode`` as
    ``MaxScoreOverRotations.result`` does (tme/analyzer/aggregation.py:
    247-261), as device tensors (the keyword arguments are the scan's
    callback arguments)."""
    torch = _torch()
    from .matching_utils import apply_convolution_mode

    scores, rotations = state[0], state[1]
    if fourier_shift is not None:
        dims = tuple(range(len(fourier_shift)))
        shift = tuple(int(x) for x in fourier_shift)
        scores = torch.roll(scores, shifts=shift, dims=dims)
        rotations = torch.roll(rotations, shifts=shift, dims=dims)
    if convolution_mode is not None:
        kw = dict(s1=targetshape, s2=templateshape, convolution_mode=convolution_mode,
                  convolution_shape=convolution_shape)
        scores = apply_convolution_mode(scores, **kw)
        rotations = apply_convolution_mode(rotations, **kw)
    return scores.contiguous(), rotations.contiguous()


def rotation_table(mapping: Dict, device=None):
    """(n, 3, 3) float32 tensor of an analyzer's rotation mapping, row i
    the matrix of rotation index i: the mapping holds index -> matrix
    (``inversion_mapping``, one rotation job) or float32 matrix bytes ->
    index (tme/analyzer/aggregation.py:170-200)."""
    torch = _torch()
    if not mapping:
        return torch.zeros((0, 3, 3), dtype=torch.float32, device=device)
    first = next(iter(mapping))
    if isinstance(first, (bytes, bytearray)):
        table = np.zeros((len(mapping), 3, 3), dtype=np.float32)
        for key, idx in mapping.items():
            table[int(idx)] = np.frombuffer(key, dtype=np.float32).reshape(3, 3)
        return torch.as_tensor(table, device=device)
    mats = [torch.as_tensor(mapping[i]) for i in range(len(mapping))]
    return torch.stack([m.to(device=device if device is not None else m.device) for m in mats]).to(torch.float32)


def rotation_matrix_to_euler_zyz(R):
    """zyz Euler angles (radians, FP64) of rotation matrices (..., 3, 3) in
    the convention of scipy's ``Rotation.as_euler("zyz")`` (extrinsic,
    R = Rz(c) Ry(b) Rz(a), b in [0, pi]), which pyTME uses
    (tme/rotations.py:199-224).  At gimbal lock (sin b = 0) the third angle
    is 0, as scipy sets it."""
    torch = _torch()
    R = R.to(torch.float64)
    r22 = R[..., 2, 2]
    sb = torch.hypot(R[..., 2, 0], R[..., 2, 1])
    b = torch.atan2(sb, r22)
    a = torch.atan2(R[..., 2, 1], -R[..., 2, 0])
    c = torch.atan2(R[..., 1, 2], R[..., 0, 2])
    lock = sb < 1e-7
    # with c = 0: b = 0 gives R = Rz(a) (R00 = cos a, R10 = sin a); b = pi
    # gives R = Ry(pi) Rz(a) (R00 = -cos a, R10 = sin a)
    a_lock = torch.where(r22 > 0, torch.atan2(R[..., 1, 0], R[..., 0, 0]), torch.atan2(R[..., 1, 0], -R[..., 0, 0]))
    a = torch.where(lock, a_lock, a)
    c = torch.where(lock, torch.zeros_like(c), c)
    return torch.stack([a, b, c], -1)


def peak_poses(positions, rotation_indices, rotation_table, template_shape):
    """Start poses of peaks at ``positions`` (n, 3) of the result volume:
    the template-corner translations positions - T/2 (the caller's
    convention, scripts/postprocess.py:598-606) and the zyz angles (radians)
    of ``rotation_table[rotation_indices[position]]``.  Returns
    ``(translations, angles, valid)``; peaks whose voxel was never updated
    (index -1) are not valid."""
    torch = _torch()
    pos = positions.to(torch.int64)
    idx = rotation_indices[tuple(pos.T)].to(torch.int64)
    valid = idx >= 0
    mats = rotation_table[torch.clamp(idx, min=0)]
    angles = rotation_matrix_to_euler_zyz(mats)
    half = torch.as_tensor(np.divide(template_shape, 2), dtype=torch.float64, device=pos.device)
    return pos.to(torch.float64) - half, angles, valid


def call_and_refine(state, callback_args: Dict, rotation_table, target, template, template_mask=None,
                    num_peaks: int = 1000, min_distance: int = 5, min_boundary_distance: Optional[int] = None,
                    min_score: Optional[float] = None, steps: int = 50, bounds_translation: Optional[float] = 5.0,
                    local=None, **refine_kwargs) -> Tuple[np.ndarray, np.ndarray, np.ndarray, Dict]:
    """Peaks of a scan's ``MaxScoreOverRotations`` state, refined.

    ``state`` is the analyzer state on the device, ``callback_args`` the
    scan's callback arguments (fourier shift, convolution mode, shapes),
    ``rotation_table`` the (n_rot, 3, 3) device tensor of the scanned
    rotations in index order (:func:`rotation_table` of the analyzer's
    mapping), ``target`` the tomogram (a device tensor is used in place).
    ``min_boundary_distance`` defaults to half the template, so every
    candidate's template box lies inside the target.  Returns the refined
    template-corner translations, zyz angles (degrees) and scores of the
    valid peaks (host arrays) and a dict with the called peaks and their
    scan scores.  ``local`` is an existing :class:`~pytme_amd.refine.LocalFLC`
    of this target and template (created when omitted).
    """
    torch = _torch()
    from .analyzer import PeakCallerMaximumFilter
    from .refine import LocalFLC, refine_peaks

    scores, rot_idx = device_result(state, **callback_args)
    if min_boundary_distance is None:
        min_boundary_distance = int(max(np.shape(template))) // 2
    pc = PeakCallerMaximumFilter(shape=tuple(scores.shape), num_peaks=num_peaks, min_distance=min_distance,
                                 min_boundary_distance=min_boundary_distance, min_score=min_score)
    peaks = pc(pc.init_state(), scores, rotation_matrix=torch.eye(3, device=scores.device))
    positions, _, peak_scores, _ = peaks
    called = (positions >= 0).all(dim=1)
    positions, peak_scores = positions[called], peak_scores[called]
    t0, a0, valid = peak_poses(positions, rot_idx, rotation_table.to(scores.device), np.shape(template))
    t0, a0, positions, peak_scores = t0[valid], a0[valid], positions[valid], peak_scores[valid]
    n = int(t0.shape[0])
    if n == 0:
        empty = np.zeros((0, 3))
        return empty, empty, np.zeros(0), {"positions": empty.astype(np.int64), "scan_scores": np.zeros(0)}
    if local is None or local.max_poses < min(n, 4096):
        local = LocalFLC(target, template, template_mask, max_poses=min(n, 4096), device=scores.device)
    t, a_deg, s, info = refine_peaks(local, t0, torch.rad2deg(a0), steps=steps,
                                     bounds_translation=bounds_translation, **refine_kwargs)
    info = dict(info)
    info["positions"] = positions.cpu().numpy()
    info["scan_scores"] = peak_scores.cpu().numpy()
    info["start_translations"] = t0.cpu().numpy()
    info["start_angles"] = np.rad2deg(a0.cpu().numpy())
    return t, a_deg, s, info
