"""Direct (non-FFT) restatements used to pin the FFT oracle on tiny inputs.

TEST INFRASTRUCTURE ONLY.  Pure NumPy loops over output positions; sizes are
kept small (<= 16^3 targets).

Coordinate convention (pinned by the reference integration case,
tests/test_matching_exhaustive.py:16-97): in ``same`` mode the score at p
correlates the target with the template whose index T//2 sits on p, with
the target zero-extended to the FFT shape and indices wrapping modulo it.
"""

import numpy as np
from scipy.ndimage import affine_transform

from .scan import next_fast_len, pull_matrix


def trilinear(arr, rotation):
    """Order-1 rotation by explicit corner weights (constant mode)."""
    arr = np.asarray(arr, dtype=np.float64)
    m = pull_matrix(rotation, (np.asarray(arr.shape) - 1) / 2)
    out = np.zeros_like(arr)
    for o in np.ndindex(*arr.shape):
        c = m[:3, :3] @ np.asarray(o, dtype=np.float64) + m[:3, 3]
        if np.any(c < 0) or np.any(c > np.asarray(arr.shape) - 1):
            continue
        i0 = np.floor(c).astype(int)
        t = c - i0
        acc = 0.0
        for corner in np.ndindex(2, 2, 2):
            idx = np.minimum(i0 + np.asarray(corner), np.asarray(arr.shape) - 1)
            w = np.prod([t[k] if corner[k] else 1 - t[k] for k in range(3)])
            acc += w * arr[tuple(idx)]
        out[o] = acc
    return out


def local_correlation(target, template, mask, rotation, score="FLC", eps=None):
    """Brute-force FLC / FLCSphericalMask / CC scores in ``same`` layout.

    template/mask are the user's (not reversed); the template is rotated
    about its centre, normalised inside the (rotated for FLC, unrotated for
    FLCSphericalMask) mask and slid over the zero-extended, circularly
    indexed target.
    """
    eps = float(np.finfo(np.float32).eps) if eps is None else eps
    target = np.asarray(target, dtype=np.float64)
    template = np.asarray(template, dtype=np.float64)
    mask = np.ones_like(template) if mask is None else np.asarray(mask, dtype=np.float64)
    fast = [next_fast_len(max(a, b)) for a, b in zip(target.shape, template.shape)]
    f = np.zeros(fast)
    f[tuple(slice(0, n) for n in target.shape)] = target
    t = affine_transform(template, pull_matrix(rotation, (np.asarray(template.shape) - 1) / 2), order=1,
                         mode="constant")
    if score == "FLC":
        m = affine_transform(mask, pull_matrix(rotation, (np.asarray(mask.shape) - 1) / 2), order=1,
                             mode="constant")
    else:
        m = mask
    n = m.sum()
    if score in ("FLC", "FLCSphericalMask"):
        mu = (t * m).sum() / n
        sd = np.sqrt(max((t * t * m).sum() / n - mu * mu, 0))
        t = (t - mu) / sd * m
    half = np.asarray(template.shape) // 2
    out = np.zeros(target.shape)
    for p in np.ndindex(*target.shape):
        idx = tuple(
            (np.arange(template.shape[k]) + p[k] - half[k]) % fast[k] for k in range(target.ndim)
        )
        win = f[np.ix_(*idx)]
        x = (win * t).sum()
        if score == "CC":
            out[p] = x
            continue
        s1 = (win * m).sum() / n
        s2 = (win * win * m).sum() / n
        sd = np.sqrt(max(s2 - s1 * s1, 0))
        sd = 1.0 if sd < eps else sd
        out[p] = x / (n * sd)
    return out
