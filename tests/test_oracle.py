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
