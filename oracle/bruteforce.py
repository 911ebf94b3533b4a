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


def local_correlation(target, template, mask, rotation, score="FLC", eps=None, template_kernel=None):
    """Brute-force FLC / FLCSphericalMask / CC scores in ``same`` layout.

    template/mask are the user's (not reversed); the template is rotated
    about its centre, optionally convolved circularly at template size with
    ``template_kernel`` (the template filter, applied after the rotation and
    before the normalisation, tme/matching_scores.py:610-613, 741-744),
    normalised inside the (rotated for FLC, unrotated for FLCSphericalMask)
    mask and slid over the zero-extended, circularly indexed target.
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
    if template_kernel is not None:
        t = circular_convolve(t, template_kernel)
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


def circular_convolve(arr, kernel):
    """Circular convolution at ``arr.shape`` by direct summation over the
    kernel's nonzero taps: out[p] = sum_j kernel[j] * arr[(p - j) mod n].

    Independent of the convolution theorem; pins the Fourier-space filter
    step of the oracle (tme/matching_scores.py:121-135,
    tme/matching_exhaustive.py:106-117).
    """
    arr = np.asarray(arr, dtype=np.float64)
    kernel = np.asarray(kernel, dtype=np.float64)
    if kernel.shape != arr.shape:
        raise ValueError("kernel and array must share a shape")
    out = np.zeros_like(arr)
    for j in np.argwhere(kernel != 0):
        out += kernel[tuple(j)] * np.roll(arr, tuple(int(x) for x in j), axis=tuple(range(arr.ndim)))
    return out


def flip_kernel(kernel):
    """kernel[-j mod n]: the kernel seen from the reversed template frame.

    The reference filters the reversed, rotated template
    (tme/matching_data.py:698-715, tme/matching_scores.py:610-612), and
    rev(a) (*) k = rev(a (*) flip(k)) for circular convolution (*), so in the
    un-reversed frame of ``local_correlation`` the template is convolved
    with the flipped kernel.
    """
    kernel = np.asarray(kernel)
    return np.roll(kernel[(slice(None, None, -1),) * kernel.ndim], 1, axis=tuple(range(kernel.ndim)))


def filtered_target(target, kernel, fast_shape):
    """Target filter of the setup: zero-extend to ``fast_shape``, convolve
    circularly with ``kernel`` (fast_shape), crop back to the target shape
    (tme/matching_exhaustive.py:106-117)."""
    f = np.zeros(fast_shape)
    f[tuple(slice(0, n) for n in np.shape(target))] = target
    return circular_convolve(f, kernel)[tuple(slice(0, n) for n in np.shape(target))]


def step_wedge_support(shape, angles, opening_axis, tilt_axis, frequency_cutoff=0.5, margin=1e-5):
    """Geometric restatement of the binary step wedge
    (tme/filters/wedge.py:298-364, 414-507 with weight_wedge=False).

    The reference sums, per tilt, the order-1 rotation of a plane holding a
    unit line at row P0//2 (plane P0 x P1 = shape[opening] x
    (shape[tilt] rounded up to odd)), pulled through
    in = R^-1 (o - c) + c about c = (P - 1)/2.  Under bilinear interpolation
    with constant (zero) borders that image is positive exactly where the
    pulled coordinate (y, x) lies inside the plane and |y - P0//2| < 1.  The
    union over tilts is cropped to the first shape[tilt] columns (centred
    crop of P1 -> shape[tilt]), spread along the remaining axis, limited to
    |k| <= cutoff, moved to the FFT origin (ifftshift).

    Returns ``(support, edge)`` in FFT order; ``edge`` marks voxels whose
    distance to some tilted line is within ``margin`` of 1, where the
    float32 rotation matrices of the reference may decide either way.
    """
    shape = tuple(int(x) for x in shape)
    n_open, n_tilt = shape[opening_axis], shape[tilt_axis]
    p0, p1 = n_open, n_tilt + (1 - n_tilt % 2)
    aspect = p0 / p1
    ang = np.radians(np.degrees(np.arctan(np.tan(np.radians(np.asarray(angles, np.float64))) * aspect)))
    row = p0 // 2
    c = np.array([(p0 - 1) / 2, (p1 - 1) / 2])
    oy, ox = np.meshgrid(np.arange(p0, dtype=np.float64), np.arange(p1, dtype=np.float64), indexing="ij")
    sup = np.zeros((p0, p1), bool)
    edge = np.zeros((p0, p1), bool)
    for a in ang:
        # scipy Rotation.from_euler("z", a) restricted to the first two axes
        r = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
        rinv = r.T
        y = rinv[0, 0] * (oy - c[0]) + rinv[0, 1] * (ox - c[1]) + c[0]
        x = rinv[1, 0] * (oy - c[0]) + rinv[1, 1] * (ox - c[1]) + c[1]
        d = np.abs(y - row)
        sup |= (d < 1) & (y >= 0) & (y <= p0 - 1) & (x >= 0) & (x <= p1 - 1)
        if a != 0:  # the identity rotation is exact in float32 too
            edge |= np.abs(d - 1) < margin
    start = (p1 - n_tilt) // 2
    out = []
    for a in (sup, edge):
        a = a[:, start : start + n_tilt]
        if opening_axis > tilt_axis:
            a = a.T
        dims = tuple(x if i in (opening_axis, tilt_axis) else 1 for i, x in enumerate(shape))
        a = np.tile(a.reshape(dims), tuple(s // d for s, d in zip(shape, dims)))
        grids = np.meshgrid(*[(np.arange(n) - n // 2) / n for n in shape], indexing="ij")
        a &= np.sqrt(sum(g * g for g in grids)) <= frequency_cutoff
        out.append(np.roll(a, tuple(n // 2 + n % 2 for n in shape), axis=tuple(range(len(shape)))))
    return tuple(out)
