"""HIP kernels against FP64 NumPy/SciPy on the GPU (through the C-ABI)."""

import numpy as np
import pytest

from oracle import scan as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    from pytme_amd.backends import backend

    backend.change_backend("hip", device="cuda")
    return backend


def _t(be, a):
    return be.to_backend_array(np.ascontiguousarray(a))


FFT_SHAPES = [
    (8,), (200,), (1024,), (45,), (7, 11), (64, 128), (24, 32, 48), (13, 10, 9), (48, 48, 48),
    (50, 60, 26), (128, 128, 128), (64, 64, 200),
]


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_rfftn_matches_numpy(be, shape):
    rng = np.random.default_rng(len(shape) * 1000 + shape[0])
    x = rng.standard_normal(shape).astype(np.float32)
    got = be.to_numpy_array(be.rfftn(_t(be, x)))
    ref = np.fft.rfftn(x.astype(np.float64))
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"rfftn {shape}: rel err {rel:.2e}", flush=True)
    assert rel < 2e-6


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_irfftn_matches_numpy(be, shape):
    rng = np.random.default_rng(7 + shape[-1])
    x = rng.standard_normal(shape)
    spec = np.fft.rfftn(x) * (1 + 0.1 * rng.standard_normal(shape[:-1] + (shape[-1] // 2 + 1,)))
    spec = spec.astype(np.complex64)
    got = be.to_numpy_array(be.irfftn(_t(be, spec), s=shape))
    ref = np.fft.irfftn(spec.astype(np.complex128), s=shape)
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    assert rel < 2e-6


def test_fft_roundtrip_full_size(be):
    # size-independent property at the tomogram size: irfftn(rfftn(x)) == x
    import torch

    x = torch.randn((512, 512, 200), device="cuda", dtype=torch.float32)
    rf, irf = be.build_fft((512, 512, 200), (512, 512, 101), inv_output_shape=(512, 512, 200))
    spec = torch.empty((512, 512, 101), dtype=torch.complex64, device="cuda")
    out = torch.empty_like(x)
    rf(x, spec)
    # Parseval on the half spectrum
    w = torch.ones(101, device="cuda")
    w[1:100] = 2
    e_spec = (spec.abs().double() ** 2 * w).sum() / x.numel()
    e_x = (x.double() ** 2).sum()
    assert abs(float(e_spec / e_x) - 1) < 1e-5
    irf(spec, out)
    assert float((out - x).abs().max()) < 5e-5


def test_unsupported_fft_length(be):
    with pytest.raises(ValueError):
        be.rfftn(_t(be, np.zeros((17 * 19,), np.float32)))


@pytest.mark.parametrize("tshape", [(9, 10, 11), (32, 32, 32), (48, 48, 48)])
@pytest.mark.parametrize("seed", [0, 1])
def test_rotation_matches_scipy(be, tshape, seed):
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(seed)
    t = rng.standard_normal(tshape).astype(np.float32)
    m = (rng.random(tshape) > 0.5).astype(np.float32)
    r = Rotation.random(random_state=seed + 5).as_matrix().astype(np.float32)
    big = tuple(s + 7 for s in tshape)
    out = be.zeros(big, be._float_dtype)
    outm = be.zeros(big, be._float_dtype)
    be.rigid_transform(_t(be, t), _t(be, r), arr_mask=_t(be, m), out=out, out_mask=outm,
                       use_geometric_center=True, order=1)
    got, gotm = be.to_numpy_array(out), be.to_numpy_array(outm)
    ref = orc.rotate(t, r).astype(np.float32)
    refm = orc.rotate(m, r).astype(np.float32)
    box = tuple(slice(0, s) for s in tshape)
    assert np.abs(got[box] - ref).max() <= 1e-6
    assert np.abs(gotm[box] - refm).max() <= 1e-6
    rest = np.ones(big, bool)
    rest[box] = False
    assert np.all(got[rest] == 0) and np.all(gotm[rest] == 0)


def test_norm_scores_and_max(be):
    rng = np.random.default_rng(3)
    n = 1 << 20
    x = rng.standard_normal(n).astype(np.float32)
    s1 = rng.standard_normal(n).astype(np.float32) * 5
    s2 = (s1.astype(np.float64) ** 2 / 77.0 + rng.random(n) * 30).astype(np.float32)
    s2[:100] = (s1[:100].astype(np.float64) ** 2 / 77.0).astype(np.float32)  # sigma ~ 0 branch
    out = be.zeros((n,), be._float_dtype)
    be.norm_scores(_t(be, x), _t(be, s2), _t(be, s1), 77.0, be.eps(be._float_dtype), out)
    got = be.to_numpy_array(out)
    # the reference's float32 operation order (npfftw_backend.py:537-587)
    n32, eps32 = np.float32(77.0), np.float32(np.finfo(np.float32).eps)
    mean2 = np.square(s1 / n32)
    sd = np.sqrt(np.maximum(s2 / n32 - mean2, np.float32(0)))
    sd[sd < eps32] = 1
    ref32 = x / (sd * n32)
    assert np.allclose(got, ref32, rtol=2e-6, atol=0)
    # and against FP64 where the variance is well conditioned
    ref = orc.norm_scores(x.astype(np.float64), s2.astype(np.float64), s1.astype(np.float64), 77.0,
                          float(np.finfo(np.float32).eps))
    assert np.abs(got[100:] - ref[100:]).max() < 1e-5 * max(1, np.abs(ref[100:]).max())

    best = _t(be, np.zeros(n, np.float32))
    rot = _t(be, np.full(n, -1, np.int32))
    be.max_score_over_rotations(_t(be, x), best, rot, 0)
    be.max_score_over_rotations(_t(be, 2 * x), best, rot, 1)
    be.max_score_over_rotations(_t(be, 2 * x), best, rot, 2)  # tie: keeps 1
    b, r = be.to_numpy_array(best), be.to_numpy_array(rot)
    assert np.array_equal(b, np.maximum(np.maximum(x, 2 * x), 0))
    assert np.array_equal(r, np.where(x > 0, 1, -1))


def test_sum_and_normalize(be):
    rng = np.random.default_rng(4)
    t = rng.standard_normal((37, 41, 23)).astype(np.float32) * 3 + 1
    m = (rng.random(t.shape) > 0.4).astype(np.float32)
    n = be.sum(_t(be, m))
    assert float(n) == float(m.astype(np.float64).sum())
    tt = _t(be, t)
    from pytme_amd.matching_utils import normalize_template

    normalize_template(tt, _t(be, m), n)
    ref = orc.normalize_template(t.astype(np.float64), m.astype(np.float64), m.sum())
    assert np.abs(be.to_numpy_array(tt) - ref).max() < 2e-6
    # deterministic: same bits on repeat
    assert float(be.sum(_t(be, t))) == float(be.sum(_t(be, t)))
