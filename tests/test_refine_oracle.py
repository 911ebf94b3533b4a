"""Pins of the local FLC refinement oracle (oracle/refine.py) on the CPU."""

import numpy as np
import pytest

from oracle import refine as orf
from oracle import scan as orc


def _volumes(seed, T=7, shape=(15, 14, 13)):
    rng = np.random.default_rng(seed)
    template = rng.normal(size=(T,) * 3)
    mask = np.zeros((T,) * 3)
    c = T // 2
    zz, yy, xx = np.indices((T,) * 3)
    mask[(zz - c) ** 2 + (yy - c) ** 2 + (xx - c) ** 2 <= (c - 0.5) ** 2] = 1
    target = rng.normal(size=shape)
    return target, template, mask


def _trilinear_loop(arr, p):
    """Explicit order-1 interpolation with zeros outside [0, n-1]."""
    n = arr.shape
    if any(pi < 0 or pi > ni - 1 for pi, ni in zip(p, n)):
        return 0.0
    i0 = [int(np.floor(pi)) for pi in p]
    fr = [pi - ii for pi, ii in zip(p, i0)]
    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                idx = [min(i0[0] + dz, n[0] - 1), min(i0[1] + dy, n[1] - 1), min(i0[2] + dx, n[2] - 1)]
                w = (fr[0] if dz else 1 - fr[0]) * (fr[1] if dy else 1 - fr[1]) * (fr[2] if dx else 1 - fr[2])
                out += w * arr[tuple(idx)]
    return out


def _score_loops(target, template_std, mask, x, eps=orf.EPS32, subvoxel="reference"):
    """Voxel-by-voxel restatement of tme/matching_optimization.py:198-548
    (subvoxel="rigid": the template centre at t + c)."""
    R = orf.euler_to_matrix(x[3:]).astype(np.float64)
    t = np.asarray(x[:3], dtype=np.float64)
    v = t.astype(int)
    s = t - v
    if subvoxel == "rigid":
        s = -(R.T @ s)
    T = template_std.shape
    c = np.floor(np.asarray(T) / 2)
    g = np.zeros(T)
    m = np.zeros(T)
    for o in np.ndindex(*T):
        q = R.T @ (np.asarray(o) - c) + c + s
        g[o] = _trilinear_loop(template_std, q)
        m[o] = _trilinear_loop(mask, q)
    n = m.sum()
    mean = (g * m).sum() / n
    sd = np.sqrt(max((g * g * m).sum() / n - mean * mean, 0))
    g = (g - mean) / sd * m
    ov = sq = ss = 0.0
    for o in np.ndindex(*T):
        p = np.asarray(o) + v
        if all(0 <= pi < ni for pi, ni in zip(p, target.shape)):
            f = target[tuple(p)]
            ov += g[o] * f
            sq += f * f * m[o]
            ss += f * m[o]
    den = np.sqrt(max(sq / n - (ss / n) ** 2, 0))
    return 0.0 if den < eps else ov / (den * n)


@pytest.mark.parametrize("x", [(4.3, 3.7, 2.2, 10.0, 20.0, -15.0), (-1.6, 9.4, 0.5, -70.0, 35.0, 120.0),
                               (3.0, 3.0, 3.0, 0.0, 0.0, 0.0)])
@pytest.mark.parametrize("subvoxel", ["reference", "rigid"])
def test_score_matches_voxel_loops(x, subvoxel):
    target, template, mask = _volumes(1)
    ts = orf.prepare_template(template, mask)
    ref = _score_loops(target, ts, mask, x, subvoxel=subvoxel)
    assert abs(orf.flc_local_score(target, ts, mask, x, subvoxel=subvoxel) - ref) < 1e-12
    assert abs(orf.flc_local_score_direct(target, ts, mask, x, subvoxel=subvoxel) - ref) < 1e-12


@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (90.0, 0.0, 0.0), (0.0, 90.0, 90.0)])
def test_self_match_scores_one(angles):
    """A positive affine copy of the (rotated) template under its box scores 1."""
    target, template, mask = _volumes(2, T=9, shape=(20, 21, 22))
    ts = orf.prepare_template(template, mask)
    v = np.array([5, 6, 7])
    R = orf.euler_to_matrix(angles)
    placed = orf.sample(template, R, np.zeros(3))
    target[v[0]:v[0] + 9, v[1]:v[1] + 9, v[2]:v[2] + 9] = 3.0 * placed + 1.5
    s = orf.flc_local_score(target, ts, mask, (*v.astype(float), *angles))
    assert abs(s - 1.0) < 1e-9


def test_matches_exhaustive_flc_at_integer_poses():
    """At integer translations and R = I the local score is the exhaustive FLC
    volume of the scan oracle at the template centre (interior positions)."""
    target, template, mask = _volumes(3, T=7, shape=(16, 15, 14))
    raw_mask = mask.copy()
    scores, _ = orc.scan("FLC", target, template, np.eye(3, dtype=np.float32)[None], template_mask=raw_mask,
                         score_threshold=-10.0)
    ts = orf.prepare_template(template, mask)
    c = 7 // 2
    checked = 0
    for p in [(5, 6, 5), (4, 4, 4), (9, 8, 7), (6, 10, 8)]:
        v = np.asarray(p) - c
        if np.any(v < 0) or np.any(v + 7 > np.asarray(target.shape)):
            continue
        local = orf.flc_local_score(target, ts, mask, (*v.astype(float), 0.0, 0.0, 0.0))
        assert abs(local - scores[p]) < 1e-9, (p, local, scores[p])
        checked += 1
    assert checked >= 3


def test_torch_sums_match_numpy():
    torch = pytest.importorskip("torch")
    euler_zyz, _, sums, _ = orf.torch_reference()
    target, template, mask = _volumes(4)
    ts = orf.prepare_template(template, mask)
    x = np.array([4.3, 3.7, 2.2, 10.0, 20.0, -15.0])
    v, s = orf.split_translation(x[:3])
    R = euler_zyz(torch.tensor(np.deg2rad(x[3:])))
    for subvoxel in ("rigid", "reference"):
        got = sums(target, ts, mask, v, torch.tensor(s), R, subvoxel).numpy()
        # the numpy path uses the float32 matrix of tme/rotations.py; compare with the same R
        want = orf.local_sums(target, ts, mask, x[:3], R.numpy(), subvoxel)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-10)
    assert np.allclose(R.numpy(), orf.euler_to_matrix(x[3:]), atol=1e-6)


@pytest.mark.parametrize("h_t, h_r, tol", [(1e-6, 1e-6, 1e-5), (1e-2, 1e-3, 5e-2)])
def test_torch_gradient_matches_finite_differences(h_t, h_r, tol):
    """Central differences inside one voxel cell.  The trilinear interpolant
    is piecewise linear, so a difference over +-h averages the slopes of the
    cells the samples cross: at h = 1e-6 it is the derivative (tight), at the
    refinement step sizes h = 1e-2 voxel / 1e-3 rad it agrees to a few per
    cent on a smooth template."""
    torch = pytest.importorskip("torch")
    from scipy.ndimage import gaussian_filter

    euler_zyz, _, sums, score = orf.torch_reference()
    target, template, mask = _volumes(5, T=9, shape=(18, 18, 18))
    template = gaussian_filter(template, 1.2)
    ts = orf.prepare_template(template, mask)
    v = np.array([4, 5, 4])
    s0 = np.array([0.31, 0.62, 0.47])
    a0 = np.deg2rad([12.0, 25.0, -8.0])

    def f(s, a):
        return float(score(sums(target, ts, mask, v, torch.tensor(s), euler_zyz(torch.tensor(a)))))

    st = torch.tensor(s0, requires_grad=True)
    at = torch.tensor(a0, requires_grad=True)
    val = score(sums(target, ts, mask, v, st, euler_zyz(at)))
    val.backward()
    scale = max(float(st.grad.abs().max()), float(at.grad.abs().max()))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h_t
        fd = (f(s0 + e, a0) - f(s0 - e, a0)) / (2 * h_t)
        assert abs(fd - float(st.grad[k])) <= tol * scale, ("t", k, fd, float(st.grad[k]))
        e[k] = h_r
        fd = (f(s0, a0 + e) - f(s0, a0 - e)) / (2 * h_r)
        assert abs(fd - float(at.grad[k])) <= tol * scale, ("r", k, fd, float(at.grad[k]))


def test_rigid_model_is_continuous_and_reference_jumps():
    """Across an integer of t the rigid model's score is continuous (the mask,
    radius 2 in a 9^3 box, stays inside the box); the reference's sampling
    puts the template at v - s, so its score just below an integer is the
    score two voxels lower, not the one at the integer (the documented
    defect)."""
    target, template, _ = _volumes(6, T=9, shape=(20, 20, 20))
    zz, yy, xx = np.indices((9,) * 3)
    mask = (((zz - 4) ** 2 + (yy - 4) ** 2 + (xx - 4) ** 2) <= 4).astype(float)
    ts = orf.prepare_template(template, mask)
    ang = (0.0, 0.0, 0.0)
    below = orf.flc_local_score(target, ts, mask, (5.0 - 1e-9, 6.0, 4.0, *ang))
    above = orf.flc_local_score(target, ts, mask, (5.0, 6.0, 4.0, *ang))
    assert abs(below - above) < 1e-6
    ang = (20.0, 35.0, -10.0)
    below = orf.flc_local_score(target, ts, mask, (5.0 - 1e-9, 6.3, 4.6, *ang))
    above = orf.flc_local_score(target, ts, mask, (5.0, 6.3, 4.6, *ang))
    assert abs(below - above) < 1e-6
    # reference sampling at R = I: t = 5 - 1e-9 is v = 4, s ~ 1, i.e. the template at 3
    r_below = orf.flc_local_score(target, ts, mask, (5.0 - 1e-9, 6.0, 4.0, 0, 0, 0), subvoxel="reference")
    r_three = orf.flc_local_score(target, ts, mask, (3.0, 6.0, 4.0, 0, 0, 0), subvoxel="reference")
    r_five = orf.flc_local_score(target, ts, mask, (5.0, 6.0, 4.0, 0, 0, 0), subvoxel="reference")
    assert abs(r_below - r_three) < 1e-6
    assert abs(r_below - r_five) > 1e-3
