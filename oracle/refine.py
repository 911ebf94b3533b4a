"""CPU oracle of the local FLC refinement score (test infrastructure only).

An FP64 restatement of the reference's derivative-free refinement objective
``matching_optimization.FLC`` (tme/matching_optimization.py:488-548) with the
pose handling of ``_MatchDensityToDensity.score`` / ``rotate_array``
(tme/matching_optimization.py:123-238) and the caller's conventions
(scripts/postprocess.py:588-613).  One deliberate difference, as the survey
of the reference records: ``rotate_array`` writes the rotated mask into the
template buffer (tme/matching_optimization.py:162-163), which zeroes the
rotated mask and makes the score undefined; here the rotated mask goes into
the mask buffer, which is what the code intends.

A second defect of the reference's pose handling is kept as an option
(``subvoxel="reference"``) and corrected by default (``subvoxel="rigid"``):
the reference samples the template at R^T (o - c) + c + s with s = t - v,
which moves the template content by -s while the integer part moves it by
+v, so the template sits at v - s: within each voxel the pose moves
backwards and it jumps by two voxels at every integer of t (and the shift is
not rotated with the template).  The rigid model places the rotated template
centre at t + c, i.e. samples at R^T (o - c - s) + c, which is continuous in
t and is the objective the refinement optimises.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` and nothing in ``pytme_amd``.

Pose x = (t_z, t_y, t_x, a, b, c): translation in voxels (the template's
corner sits at the integer part v = trunc(t) in the target) and zyz Euler
angles in degrees (scipy ``from_euler("zyz")``, cast to float32 as
tme/rotations.py:171-196).  The template is sampled at
q = R^T (o - c) + c + (t - v), c = floor(T / 2), with order-1 ``constant``
interpolation (scipy.ndimage.map_coordinates), the rotated template is
standardised inside the rotated mask, and

    score = sum(t_R * f) / (n * sd_f),  sd_f = sqrt(max(E_m[f^2] - E_m[f]^2, 0)),

over the overlap of the template box with the target (voxels of the box
outside the target count in n and in the standardisation, not in the sums
with f).  sd_f < eps gives 0 (tme/matching_optimization.py:542-544).

Pinning (tests/test_refine_oracle.py):
  * a loop-by-loop brute-force restatement on small volumes;
  * the self-match closed form: an exact copy of the template, at any
    positive affine intensity, scores 1 at its pose;
  * at integer translations and the identity rotation the local score equals
    the exhaustive FLC score volume of ``oracle.scan`` at the template centre
    (two independent definitions of the same correlation);
  * the torch FP64 autograd gradient equals central finite differences.
"""

from typing import Sequence, Tuple

import numpy as np
from scipy.ndimage import map_coordinates
from scipy.spatial.transform import Rotation

EPS32 = float(np.finfo(np.float32).eps)


def euler_to_matrix(angles_deg: Sequence[float]) -> np.ndarray:
    """tme/rotations.py:171-196: zyz Euler angles (degrees) -> float32 R."""
    return Rotation.from_euler("zyz", np.asarray(angles_deg, dtype=np.float64), degrees=True).as_matrix().astype(
        np.float32
    )


def split_translation(t: Sequence[float]) -> Tuple[np.ndarray, np.ndarray]:
    """tme/matching_optimization.py:212-213: v = t.astype(int) (toward zero), s = t - v."""
    t = np.asarray(t, dtype=np.float64)
    v = t.astype(np.int64)
    return v, t - v


def sample(arr: np.ndarray, R: np.ndarray, shift: Sequence[float]) -> np.ndarray:
    """tme/matching_optimization.py:123-170 (rotate_array): arr at
    R^T (o - c) + c + shift, c = floor(shape / 2), order 1, mode constant."""
    arr = np.asarray(arr, dtype=np.float64)
    c = np.floor(np.asarray(arr.shape) / 2)[:, None]
    grid = np.indices(arr.shape, dtype=np.float64).reshape(arr.ndim, -1) - c
    pos = np.asarray(R, dtype=np.float64).T @ grid + (np.asarray(shift, dtype=np.float64)[:, None] + c)
    return map_coordinates(arr, pos, order=1, mode="constant", cval=0.0).reshape(arr.shape)


def prepare_template(template: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """tme/matching_optimization.py:515-520 (_post_init): standardise the
    template inside the unrotated mask once."""
    t = np.asarray(template, dtype=np.float64)
    m = np.asarray(mask, dtype=np.float64)
    n = m.sum()
    mean = np.sum(t * m) / n
    sd = np.sqrt(max(np.sum(t * t * m) / n - mean * mean, 0.0))
    return (t - mean) / sd * m


def target_window(target: np.ndarray, v: Sequence[int], shape: Sequence[int]) -> np.ndarray:
    """tme/matching_optimization.py:215-238: the target under the template box
    with corner v; voxels of the box outside the target are 0 (they are not
    part of the overlap sums)."""
    out = np.zeros(shape, dtype=np.float64)
    src, dst = [], []
    for vi, ni, ti in zip(v, target.shape, shape):
        lo, hi = max(vi, 0), min(vi + ti, ni)
        if hi <= lo:
            return out
        src.append(slice(lo, hi))
        dst.append(slice(lo - vi, hi - vi))
    out[tuple(dst)] = target[tuple(src)]
    return out


def sampling_shift(R, s, subvoxel: str = "rigid") -> np.ndarray:
    """The shift added to R^T (o - c) + c: -R^T s for the rigid model, s for
    the reference's sampling (tme/matching_optimization.py:150-157)."""
    s = np.asarray(s, dtype=np.float64)
    if subvoxel == "rigid":
        return -(np.asarray(R, dtype=np.float64).T @ s)
    if subvoxel == "reference":
        return s
    raise ValueError("subvoxel is 'rigid' or 'reference'")


def local_sums(target, template_std, mask, translation, R, subvoxel: str = "rigid") -> np.ndarray:
    """The six per-pose moments the score is assembled from:
    (n, sum g m, sum g^2 m, sum g m f, sum m f, sum m f^2) with g, m the
    rotated template and mask and f the target window."""
    v, s = split_translation(translation)
    shift = sampling_shift(R, s, subvoxel)
    g = sample(template_std, R, shift)
    m = sample(mask, R, shift)
    f = target_window(np.asarray(target, dtype=np.float64), v, template_std.shape)
    return np.array([m.sum(), (g * m).sum(), (g * g * m).sum(), (g * m * f).sum(), (m * f).sum(), (m * f * f).sum()])


def score_from_sums(S, eps: float = EPS32) -> float:
    """tme/matching_optimization.py:522-548 written in the six moments:
    the rotated template is standardised inside the rotated mask
    (normalize_template, tme/matching_utils.py:61-99), then
    overlap / (sd_f * n)."""
    n, a, b, c, d, e = (float(x) for x in S)
    mu = a / n
    sd_t = np.sqrt(max(b / n - mu * mu, 0.0))
    overlap = (c - mu * d) / sd_t
    sd_f = np.sqrt(max(e / n - (d / n) ** 2, 0.0))
    if sd_f < eps:
        return 0.0
    return overlap / (sd_f * n)


def flc_local_score(target, template_std, mask, x: Sequence[float], eps: float = EPS32,
                    subvoxel: str = "rigid") -> float:
    """Positive FLC of pose x = (t(3), zyz degrees(3)) (the reference returns
    its negative for minimisation, tme/matching_optimization.py:546)."""
    R = euler_to_matrix(x[3:])
    return score_from_sums(local_sums(target, template_std, mask, x[:3], R, subvoxel), eps)


def flc_local_score_direct(target, template_std, mask, x: Sequence[float], eps: float = EPS32,
                           subvoxel: str = "reference") -> float:
    """The reference's own order of operations (explicit normalisation of the
    rotated template, slicing of the overlap), used to pin score_from_sums."""
    R = euler_to_matrix(x[3:])
    v, s = split_translation(x[:3])
    T = np.asarray(template_std.shape)
    shift = sampling_shift(R, s, subvoxel)
    g = sample(template_std, R, shift)
    m = sample(mask, R, shift)
    n = m.sum()
    mean = np.sum(g * m) / n
    sd = np.sqrt(max(np.sum(g * g * m) / n - mean * mean, 0.0))
    g = (g - mean) / sd * m
    starts = np.maximum(v, 0)
    stops = np.minimum(v + T, target.shape)
    ts = tuple(slice(a - vi, b - vi) for a, b, vi in zip(starts, stops, v))
    fs = tuple(slice(a, b) for a, b in zip(starts, stops))
    f = np.asarray(target, dtype=np.float64)[fs]
    overlap = np.sum(g[ts] * f)
    mr = m[ts]
    exp_sq = np.sum(f * f * mr) / n
    sq_exp = (np.sum(f * mr) / n) ** 2
    den = np.sqrt(max(exp_sq - sq_exp, 0.0))
    if den < eps:
        return 0.0
    return overlap / (den * n)


# ---------------------------------------------------------------- autograd
def torch_reference():
    """FP64 torch restatement with autograd through the interpolation
    weights (the gradient reference for the GPU backward kernel)."""
    import torch

    def euler_zyz(angles_rad):
        # scipy from_euler("zyz") (extrinsic): R = Rz(c) Ry(b) Rz(a)
        a, b, c = angles_rad[..., 0], angles_rad[..., 1], angles_rad[..., 2]

        def rz(t):
            z, o = torch.zeros_like(t), torch.ones_like(t)
            return torch.stack([torch.stack([torch.cos(t), -torch.sin(t), z], -1),
                                torch.stack([torch.sin(t), torch.cos(t), z], -1),
                                torch.stack([z, z, o], -1)], -2)

        def ry(t):
            z, o = torch.zeros_like(t), torch.ones_like(t)
            return torch.stack([torch.stack([torch.cos(t), z, torch.sin(t)], -1),
                                torch.stack([z, o, z], -1),
                                torch.stack([-torch.sin(t), z, torch.cos(t)], -1)], -2)

        return rz(c) @ ry(b) @ rz(a)

    def trilinear(arr, pos):
        # map_coordinates order 1, mode constant: 0 outside [0, n-1]
        n = torch.tensor(arr.shape, dtype=pos.dtype)
        valid = ((pos >= 0) & (pos <= (n - 1)[:, None])).all(0)
        i0 = torch.floor(pos).detach()
        fr = pos - i0
        i0 = i0.long()
        out = torch.zeros(pos.shape[1], dtype=pos.dtype)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    idx = [torch.clamp(i0[k] + d, 0, arr.shape[k] - 1) for k, d in enumerate((dz, dy, dx))]
                    w = ((fr[0] if dz else 1 - fr[0]) * (fr[1] if dy else 1 - fr[1]) * (fr[2] if dx else 1 - fr[2]))
                    out = out + w * arr[idx[0], idx[1], idx[2]]
        return torch.where(valid, out, torch.zeros_like(out))

    def sums(target, template_std, mask, v, s, R, subvoxel="rigid"):
        """Six moments for integer corner v, sub-voxel shift s (tensor, grad)
        and rotation R (tensor, grad)."""
        if subvoxel == "rigid":
            s = -(R.transpose(-1, -2) @ s)
        T = template_std.shape
        tmpl = torch.as_tensor(template_std, dtype=torch.float64)
        msk = torch.as_tensor(mask, dtype=torch.float64)
        c = torch.floor(torch.tensor(T, dtype=torch.float64) / 2)[:, None]
        grid = torch.as_tensor(np.indices(T, dtype=np.float64).reshape(3, -1)) - c
        pos = R.transpose(-1, -2) @ grid + (s[:, None] + c)
        g = trilinear(tmpl, pos)
        m = trilinear(msk, pos)
        f = torch.as_tensor(target_window(np.asarray(target, dtype=np.float64), v, T).reshape(-1))
        return torch.stack([m.sum(), (g * m).sum(), (g * g * m).sum(), (g * m * f).sum(), (m * f).sum(),
                            (m * f * f).sum()])

    def score(S, eps=EPS32):
        n, a, b, c, d, e = S
        mu = a / n
        sd_t = torch.sqrt(torch.clamp(b / n - mu * mu, min=0.0))
        overlap = (c - mu * d) / sd_t
        sd_f = torch.sqrt(torch.clamp(e / n - (d / n) ** 2, min=0.0))
        return torch.where(sd_f < eps, torch.zeros_like(sd_f), overlap / (sd_f * n))

    return euler_zyz, trilinear, sums, score
