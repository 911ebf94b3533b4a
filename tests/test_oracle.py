"""Pins of the CPU oracle (oracle/) against the reference's known answers,
closed forms and brute force.  CPU only."""

import numpy as np
import pytest
from scipy.ndimage import affine_transform, laplace
from scipy.spatial.transform import Rotation

from oracle import bruteforce as bf
from oracle import scan as orc


def _rot(seed):
    return Rotation.random(random_state=seed).as_matrix().astype(np.float32)


class TestRotation:
    def test_identity_exact(self):
        a = np.random.default_rng(0).standard_normal((6, 7, 5))
        assert np.array_equal(orc.rotate(a, np.eye(3, dtype=np.float32)), a)

    def test_90deg_is_permutation(self):
        # rotation by 90 degrees about axis 0 about the geometric centre of a
        # cube maps voxels onto voxels: exact transpose + flip
        a = np.random.default_rng(1).standard_normal((7, 7, 7))
        r = np.round(Rotation.from_euler("x", 90, degrees=True).as_matrix()).astype(np.float32)
        out = orc.rotate(a, r)
        cands = [np.rot90(a, k, axes=(1, 2)) for k in (1, 3)]
        assert min(np.abs(out - c).max() for c in cands) == 0

    def test_flip_preserves_sum(self):
        # reference tests/test_backends.py:357-393: an axis flip keeps the sum
        a = np.random.default_rng(2).random((8, 8, 8))
        r = np.diag([-1.0, -1.0, 1.0]).astype(np.float32)
        assert np.isclose(orc.rotate(a, r).sum(), a.sum(), rtol=1e-6)

    @pytest.mark.parametrize("seed", [0, 1, 2])
    def test_bruteforce_trilinear(self, seed):
        a = np.random.default_rng(seed).standard_normal((4, 4, 4))
        r = _rot(seed)
        assert np.abs(orc.rotate(a, r) - bf.trilinear(a, r)).max() < 1e-12

    def test_constant_mode_edges(self):
        # scipy order 1 / constant: coordinates beyond [0, n-1] give 0, the
        # exact edge interpolates (plain 1-D affine_transform behaviour)
        x = np.arange(1, 6, dtype=np.float64)
        out = affine_transform(x, np.array([[1.0, 0.5], [0, 1]]), order=1, mode="constant")
        assert np.allclose(out, [1.5, 2.5, 3.5, 4.5, 0.0])


class TestCorrelation:
    @pytest.mark.parametrize("score", ["CC", "FLCSphericalMask", "FLC"])
    @pytest.mark.parametrize("tshape", [(5, 5, 5), (6, 6, 6), (5, 6, 4)])
    def test_fft_vs_bruteforce(self, score, tshape):
        rng = np.random.default_rng(7)
        target = rng.standard_normal((12, 10, 9))
        template = rng.standard_normal(tshape)
        mask = (rng.random(tshape) > 0.3).astype(float)
        r = _rot(11)
        s, _ = orc.scan(score, target, template, r[None], template_mask=mask, score_threshold=-1e9)
        b = bf.local_correlation(target, template, mask, r, score=score)
        assert np.abs(s - b).max() < 1e-10

    def test_normalize_template_closed_form(self):
        # tests/test_matching_utils.py:180-189: mean 0 and std 1 inside mask
        rng = np.random.default_rng(3)
        t = rng.standard_normal((9, 9, 9)) * 3 + 2
        m = (rng.random((9, 9, 9)) > 0.4).astype(float)
        out = orc.normalize_template(t, m, m.sum())
        sel = m > 0
        assert abs(out[sel].mean()) < 1e-12 and abs(out[sel].std() - 1) < 1e-12
        assert np.all(out[~sel] == 0)

    def test_self_match_scores_one(self):
        rng = np.random.default_rng(4)
        target = rng.standard_normal((20, 20, 20))
        template = target[5:12, 6:13, 4:11].copy()
        s, _ = orc.scan("FLC", target, template, np.eye(3, dtype=np.float32)[None], score_threshold=-1)
        peak = np.unravel_index(np.argmax(s), s.shape)
        assert np.allclose(peak, (8, 9, 7)) and abs(s[peak] - 1) < 1e-9

    def test_flc_equals_flcsph_for_invariant_mask(self):
        # a mask invariant under the 90-degree rotations used gives the same
        # normalisation for FLC (rotated mask) and FLCSphericalMask (fixed)
        rng = np.random.default_rng(5)
        target = rng.standard_normal((16, 16, 16))
        template = rng.standard_normal((7, 7, 7))
        mask = np.zeros((7, 7, 7))
        mask[1:6, 1:6, 1:6] = 1
        rots = np.stack([Rotation.from_euler("z", a, degrees=True).as_matrix() for a in (0, 90, 180)])
        a, ra = orc.scan("FLC", target, template, rots, template_mask=mask, score_threshold=-1e9)
        b, rb = orc.scan("FLCSphericalMask", target, template, rots, template_mask=mask, score_threshold=-1e9)
        assert np.abs(a - b).max() < 1e-9 and np.array_equal(ra, rb)


class TestReferenceIntegrationCase:
    """tests/test_matching_exhaustive.py:16-97 of the reference."""

    def setup_method(self):
        target = np.zeros((80, 80, 80))
        target[25:31, 22:28, 12:16] = 1
        template = np.zeros((41, 41, 35))
        template[20:26, 25:31, 17:21] = 1
        self.target, self.template = target, template
        self.mask = np.ones_like(template)

    @pytest.mark.parametrize("score", ["CC", "LCC", "FLCSphericalMask", "FLC"])
    @pytest.mark.parametrize("pad_edge", [False, True])
    def test_peak(self, score, pad_edge):
        s, r = orc.scan_subsets(
            score, self.target, self.template, np.eye(3, dtype=np.float32)[None], template_mask=self.mask,
            target_splits={0: 2}, pad_target_edges=pad_edge,
        )
        peak = np.unravel_index(np.argmax(s), s.shape)
        assert np.allclose(peak, [25, 17, 12])
        expected = 1.0
        if score == "CC":
            expected = self.template.sum()
        elif score == "LCC":
            expected = (laplace(self.template) * laplace(self.template)).sum()
        assert np.allclose(s[peak], expected, rtol=0.3)


class TestAnalyzerRule:
    @pytest.mark.parametrize("threshold", [0, 1e10, -1e10])
    def test_closed_form(self, threshold):
        # tests/test_analyzer.py:171-193: max(data, 2*data, threshold)
        data = np.random.default_rng(6).random((10, 10, 10))
        best = np.full(data.shape, threshold, dtype=np.float64)
        rot = np.full(data.shape, -1, dtype=np.int32)
        orc.max_score_update(data, best, rot, 0)
        orc.max_score_update(2 * data, best, rot, 1)
        assert np.allclose(best, np.maximum(np.maximum(data, 2 * data), threshold))

    def test_tie_keeps_first(self):
        data = np.random.default_rng(8).random((5, 5))
        best = np.zeros_like(data)
        rot = np.full(data.shape, -1, dtype=np.int32)
        orc.max_score_update(data, best, rot, 0)
        orc.max_score_update(data.copy(), best, rot, 1)
        assert np.all(rot == 0)


class TestPadding:
    def test_fourier_padding_values(self):
        conv, fast, ft, shift = orc.fourier_padding((512, 512, 200), (48, 48, 48))
        assert fast == (512, 512, 200) and ft == (512, 512, 101) and shift == (-23, -23, -23)
        conv, fast, ft, shift = orc.fourier_padding((128,) * 3, (32,) * 3)
        assert shift == (-15, -15, -15)
        assert orc.fourier_padding((80,) * 3, (41, 41, 35))[3] == (-20, -20, -17)


def _kernel(shape, taps):
    k = np.zeros(shape)
    for idx, v in taps:
        k[idx] = v
    return k


# one-voxel shift, an anisotropic 3-tap blur, and an asymmetric mixture: each
# filter is the rfft half spectrum of the kernel at the array's shape
_KERNELS = {
    "shift": [((1, 0, 0), 1.0)],
    "blur": [((0, 0, 0), 0.5), ((0, 0, 1), 0.25), ((0, 0, -1), 0.25)],
    "mixed": [((0, 0, 0), 0.6), ((1, 0, 0), 0.3), ((0, -1, 2), -0.2), ((-2, 1, 0), 0.1)],
}


class TestFilterStep:
    """The oracle's Fourier-space filter steps against direct circular
    convolution (tme/matching_scores.py:121-135 template-size filter,
    tme/matching_exhaustive.py:106-117 target filter)."""

    @pytest.mark.parametrize("name", sorted(_KERNELS))
    @pytest.mark.parametrize("shape", [(6, 6, 6), (5, 6, 7), (7, 4, 5)])
    def test_template_filter_is_circular_convolution(self, name, shape):
        rng = np.random.default_rng(21)
        t = rng.standard_normal(shape)
        k = _kernel(shape, _KERNELS[name])
        filt = np.fft.rfftn(k)  # half spectrum along the last axis
        orc_ = orc.ScanOracle("CC", rng.standard_normal((12, 12, 12)), t, template_filter=filt)
        assert np.abs(orc_._apply_filter(t) - bf.circular_convolve(t, k)).max() < 1e-12

    def test_all_ones_filter_is_identity(self):
        t = np.random.default_rng(22).standard_normal((5, 6, 7))
        orc_ = orc.ScanOracle("CC", np.zeros((9, 9, 9)), t, template_filter=np.ones((5, 6, 4)))
        assert np.abs(orc_._apply_filter(t) - t).max() < 1e-12
        # a size-1 filter is the reference's "no filter" (tme/matching_scores.py:90-91)
        assert orc.ScanOracle("CC", np.zeros((9, 9, 9)), t, template_filter=np.ones(1)).filter is None

    def test_kernel_flip_is_not_a_no_op(self):
        # the reversed-frame kernel differs from the kernel for asymmetric
        # filters, so the filtered-scan pin below tells a filter applied in
        # the reversed frame from one applied in the user's frame
        rng = np.random.default_rng(23)
        shape = (5, 6, 7)
        t = rng.standard_normal(shape)
        k = _kernel(shape, _KERNELS["mixed"])
        assert np.abs(bf.circular_convolve(t, bf.flip_kernel(k)) - bf.circular_convolve(t, k)).max() > 0.1
        assert np.array_equal(bf.flip_kernel(bf.flip_kernel(k)), k)

    @pytest.mark.parametrize("name", sorted(_KERNELS))
    @pytest.mark.parametrize("tshape", [(13, 11, 9), (12, 10, 16)])
    def test_target_filter_at_fast_shape_then_crop(self, name, tshape):
        rng = np.random.default_rng(24)
        target = rng.standard_normal(tshape)
        template = rng.standard_normal((4, 5, 3))
        _, fast, _, _ = orc.fourier_padding(tshape, template.shape)
        k = _kernel(fast, _KERNELS[name])
        o = orc.ScanOracle("CC", target, template, target_filter=np.fft.rfftn(k))
        got = np.fft.irfftn(o.ft_target, s=fast)
        want = bf.filtered_target(target, k, fast)
        assert np.abs(got[tuple(slice(0, n) for n in tshape)] - want).max() < 1e-12
        # the filtered target is cropped back and zero-extended again
        outside = ~np.pad(np.ones(tshape, bool), [(0, f - n) for f, n in zip(fast, tshape)])
        assert np.all(np.abs(got[outside]) < 1e-12)

    @pytest.mark.parametrize("score", ["CC", "FLCSphericalMask", "FLC"])
    @pytest.mark.parametrize("name", ["shift", "mixed"])
    def test_filtered_scan_vs_bruteforce(self, score, name):
        rng = np.random.default_rng(25)
        target = rng.standard_normal((11, 10, 9))
        tshape = (5, 6, 4)
        template = rng.standard_normal(tshape)
        mask = (rng.random(tshape) > 0.3).astype(float)
        _, fast, _, _ = orc.fourier_padding(target.shape, tshape)
        kt = _kernel(tshape, _KERNELS[name])
        kf = _kernel(fast, _KERNELS["blur"])
        r = _rot(31)
        s, _ = orc.scan(score, target, template, r[None], template_mask=mask, template_filter=np.fft.rfftn(kt),
                        target_filter=np.fft.rfftn(kf), score_threshold=-1e9)
        b = bf.local_correlation(bf.filtered_target(target, kf, fast), template, mask, r, score=score,
                                 template_kernel=bf.flip_kernel(kt))
        assert np.abs(s - b).max() < 1e-10

    def test_filter_after_rotation_matters(self):
        # filtering before the rotation (a plausible mistake) changes the scores
        rng = np.random.default_rng(26)
        target = rng.standard_normal((11, 10, 9))
        template = rng.standard_normal((5, 6, 4))
        kt = _kernel(template.shape, _KERNELS["mixed"])
        r = _rot(32)
        s, _ = orc.scan("CC", target, template, r[None], template_filter=np.fft.rfftn(kt), score_threshold=-1e9)
        pre = bf.circular_convolve(template, kt)
        b = bf.local_correlation(target, pre, None, r, score="CC")
        assert np.abs(s - b).max() > 1e-3
