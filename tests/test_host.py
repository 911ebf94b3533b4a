"""Host-side logic and the C-ABI surface (CPU only; no GPU calls)."""

import ctypes
import os
import re

import numpy as np
import pytest
from scipy.signal import correlate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class TestBackendManager:
    def test_add_backend_rejects_non_backend(self):
        from pytme_amd.backends import BackendManager

        with pytest.raises(ValueError):
            BackendManager().add_backend("bad", dict)

    def test_unknown_backend(self):
        from pytme_amd.backends import BackendManager

        with pytest.raises(NotImplementedError):
            BackendManager().change_backend("does-not-exist")

    def test_hip_listed_and_constructible_on_cpu(self):
        from pytme_amd.backends import BackendManager, HipBackend

        m = BackendManager()
        assert "hip" in m.available_backends()
        m.add_backend("hip2", HipBackend)
        m.change_backend("hip2", device="cpu")
        assert m._backend_name == "hip2"
        z = m.zeros((2, 3), m._float_dtype)
        assert tuple(z.shape) == (2, 3)

    def test_compute_refuses_cpu_device(self, cpu_backend):
        be = cpu_backend
        x = be.zeros((4, 4, 4), be._float_dtype)
        with pytest.raises(RuntimeError):
            be.rigid_transform(x, np.eye(3, dtype=np.float32), out=be.zeros((4, 4, 4), be._float_dtype), order=1)
        with pytest.raises(RuntimeError):
            be.rfftn(x)
        with pytest.raises(RuntimeError):
            be.norm_scores(x, x, x, 1.0, 1e-7, x)

    def test_compute_convolution_shapes(self, cpu_backend):
        conv, fast, ft = cpu_backend.compute_convolution_shapes((512, 512, 200), (1, 1, 1))
        assert fast == [512, 512, 200] and ft == [512, 512, 101]
        from scipy.fft import next_fast_len

        from pytme_amd.backends.hip_backend import _next_fast_len

        for n in range(1, 600):
            assert _next_fast_len(n) == next_fast_len(n, real=True)


class TestMultiTemplateSchedule:
    def test_schedule_arguments(self, cpu_backend):
        """scan_templates refuses an unknown schedule, and the one-graph
        schedule refuses target/template splits and a backend without the
        fused GPU scan (here: the hip backend on the CPU device) before any
        setup work."""
        from pytme_amd.matching_exhaustive import scan_templates

        target = np.zeros((16, 16, 16), np.float32)
        entries = [dict(template=np.ones((4, 4, 4), np.float32), template_mask=np.ones((4, 4, 4), np.float32))]
        with pytest.raises(ValueError, match="schedule"):
            scan_templates(target, entries, score="FLC", schedule="bogus")
        with pytest.raises(ValueError, match="target_splits"):
            scan_templates(target, entries, score="FLC", schedule="graph", target_splits={0: 2})
        with pytest.raises(ValueError, match="GPU"):
            scan_templates(target, entries, score="FLC", schedule="graph")
        with pytest.raises(ValueError):
            scan_templates(target, entries, score="NoSuchScore", schedule="graph")


class TestRegistry:
    def test_register_and_duplicate(self):
        from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, register_matching_exhaustive
        from pytme_amd.memory import MATCHING_MEMORY_REGISTRY

        key = "FLC"
        setup, scoring = MATCHING_EXHAUSTIVE_REGISTER[key]
        register_matching_exhaustive("TEST_FLC", setup, scoring, MATCHING_MEMORY_REGISTRY[key])
        with pytest.raises(ValueError):
            register_matching_exhaustive(key, setup, scoring, MATCHING_MEMORY_REGISTRY[key])
        with pytest.raises(ValueError):
            register_matching_exhaustive("TEST2", setup, scoring, dict)
        del MATCHING_EXHAUSTIVE_REGISTER["TEST_FLC"]

    def test_scores_registered(self):
        from pytme_amd.matching_scores import MATCHING_EXHAUSTIVE_REGISTER

        assert set(MATCHING_EXHAUSTIVE_REGISTER) >= {"CC", "LCC", "FLC", "FLCSphericalMask"}


class TestMatchingData:
    def test_fourier_padding(self, cpu_backend):
        from oracle import scan as orc
        from pytme_amd.matching_data import MatchingData

        for tshape, Tshape in [((512, 512, 200), (48,) * 3), ((128,) * 3, (32,) * 3), ((80,) * 3, (41, 41, 35)),
                               ((20, 30, 40), (7, 8, 9))]:
            for pad in (False, True):
                assert MatchingData._fourier_padding(tshape, Tshape, pad) == orc.fourier_padding(tshape, Tshape, pad)

    def test_template_reversed(self, cpu_backend):
        from pytme_amd.matching_data import MatchingData

        t = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
        md = MatchingData(np.zeros((8, 8, 8), np.float32), t)
        assert np.array_equal(md.template, t[::-1, ::-1, ::-1])
        assert md.template_mask.shape == t.shape and np.all(md.template_mask == 1)

    def test_subset_reflect_padding(self, cpu_backend):
        from oracle import scan as orc
        from pytme_amd.matching_data import MatchingData

        a = np.random.default_rng(0).random((20, 18, 16))
        md = MatchingData(a, np.ones((5, 5, 5)))
        sl = (slice(0, 10), slice(4, 18), slice(8, 16))
        got = md.subset_array(a, sl, np.array([4, 3, 6]))
        assert np.array_equal(got, orc.subset_array(a, sl, np.array([4, 3, 6])))

    def test_split_rotations(self, cpu_backend):
        from pytme_amd.matching_data import MatchingData

        md = MatchingData(np.zeros((8, 8, 8)), np.ones((3, 3, 3)), rotations=np.tile(np.eye(3), (10, 1, 1)))
        parts = md._split_rotations_on_jobs(3)
        assert [p.shape[0] for p in parts] == [3, 3, 4]


class TestUtils:
    def test_split_shape_matches_oracle(self):
        from oracle import scan as orc
        from pytme_amd.matching_utils import split_shape

        for shape, splits in [((80, 80, 80), {0: 2}), ((17, 9, 30), {0: 3, 2: 4}), ((5, 5), {})]:
            assert split_shape(shape, splits) == orc.split_shape(shape, splits)

    @pytest.mark.parametrize("mode", ["full", "same", "valid"])
    def test_convolution_mode_shapes(self, mode):
        # reference tests/test_matching_utils.py:123-143: shapes equal scipy's
        from pytme_amd.matching_utils import apply_convolution_mode

        a, b = np.random.rand(12, 10, 9), np.random.rand(5, 4, 3)
        full = correlate(a, b, mode="full")
        out = apply_convolution_mode(full, mode, a.shape, b.shape)
        assert out.shape == correlate(a, b, mode=mode).shape

    def test_elliptical_mask(self):
        from pytme_amd.matching_utils import elliptical_mask

        m = elliptical_mask((32, 32, 32), radius=14)
        assert m[16, 16, 16] == 1 and m[16, 16, 30] == 1 and m[16, 16, 31] == 0 and m[0, 0, 0] == 0


class TestAnalyzerMerge:
    @pytest.mark.parametrize("threshold", [0, 1e10, -1e10])
    def test_merge_closed_form(self, threshold, cpu_backend):
        # reference tests/test_analyzer.py:195-221 with host results
        from pytme_amd.analyzer import MaxScoreOverRotations

        data = np.random.default_rng(1).random((20, 20, 20)).astype(np.float32)
        eye = np.eye(3, dtype=np.float32)
        r1 = (np.maximum(data, threshold).astype(np.float32), np.zeros(3, int),
              np.where(data > threshold, 0, -1).astype(np.int32), {0: eye})
        r2 = (np.maximum(2 * data, threshold).astype(np.float32), np.zeros(3, int),
              np.where(2 * data > threshold, 0, -1).astype(np.int32), {0: eye * 2})
        s, off, rot, mapping = MaxScoreOverRotations.merge([r1, r2], score_threshold=threshold)
        assert np.allclose(s, np.maximum(np.maximum(data, 2 * data), threshold))
        assert len(mapping) == 2

    def test_merge_offsets(self, cpu_backend):
        from pytme_amd.analyzer import MaxScoreOverRotations

        eye = np.eye(3, dtype=np.float32)
        a = (np.ones((4, 4, 4), np.float32), np.array([0, 0, 0]), np.zeros((4, 4, 4), np.int32), {0: eye})
        b = (2 * np.ones((4, 4, 4), np.float32), np.array([2, 0, 0]), np.zeros((4, 4, 4), np.int32), {0: eye})
        s, off, rot, mapping = MaxScoreOverRotations.merge([a, b])
        assert s.shape == (6, 4, 4) and np.all(s[:2] == 1) and np.all(s[2:] == 2)


class TestFilters:
    def test_electron_wavelength(self):
        from pytme_amd.filters.ctf import electron_wavelength

        assert abs(electron_wavelength(300e3) - 0.019687) < 1e-6

    def test_ctf_zeros_and_dc(self):
        from pytme_amd.filters.ctf import create_ctf, electron_wavelength

        n, apix, dz, A = 256, 1.0, 20000.0, 0.07
        ctf = create_ctf((n,), defocus_x=dz, sampling_rate=apix, spherical_aberration=0.0, amplitude_contrast=A)
        # with Cs = 0: chi = pi*lam*dz*k^2 + atan(A/sqrt(1-A^2)) ; DC value = -A
        assert np.isclose(ctf[n // 2], -np.sin(np.arctan(A / np.sqrt(1 - A * A))), atol=1e-6)
        lam = electron_wavelength(300e3)
        phi = np.arctan(A / np.sqrt(1 - A * A))
        k1 = np.sqrt((np.pi - phi) / (np.pi * lam * dz))  # first zero of sin(chi)
        k = (np.arange(n) - n // 2) / n
        pos = ctf[n // 2:]
        first = np.argmax(np.sign(pos[1:]) != np.sign(pos[:-1]))  # first sign change for k > 0
        assert k[n // 2 + first] <= k1 <= k[n // 2 + first + 1]
        assert np.all(np.abs(ctf) <= 1 + 1e-6)

    def test_wedge_binary_symmetric_dc(self):
        from pytme_amd.filters import WedgeReconstructed

        w = WedgeReconstructed(angles=np.arange(-60, 61, 3.0), opening_axis=2, tilt_axis=0)
        full = w(shape=(48, 48, 48), return_real_fourier=False)["data"]
        assert set(np.unique(full)) <= {0.0, 1.0}
        assert full[0, 0, 0] == 1
        # Friedel symmetry of the binary support: W(-k) == W(k)
        flipped = np.roll(full[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))
        assert np.mean(flipped == full) > 0.97
        half = w(shape=(48, 48, 48), return_real_fourier=True)["data"]
        assert half.shape == (48, 48, 25)

    def test_continuous_wedge_angles(self):
        from pytme_amd.filters import WedgeReconstructed

        w = WedgeReconstructed(angles=(-60, 60), create_continuous_wedge=True, opening_axis=2, tilt_axis=0)
        full = w(shape=(64, 64, 64), return_real_fourier=False)["data"]
        # sampled where |k_tilt| >= tan(30 deg) |k_opening|: the whole tilt
        # axis is sampled, the opening axis (beyond DC) lies in the missing wedge
        c = np.fft.fftshift(full)
        assert np.all(c[12:52, 32, 32] == 1)
        assert np.all(c[32, 32, 33:50] == 0) and c[32, 32, 32] == 1
        assert c[32 + 6, 32, 32 + 10] == 1 and c[32 + 5, 32, 32 + 10] == 0


class TestCABI:
    def test_library_exports_header(self):
        from pytme_amd import _lib

        lib = _lib.lib()
        header = open(os.path.join(ROOT, "include", "pytme_hip.h")).read()
        names = set(re.findall(r"^\s*(?:int64_t|int|const char\*)\s+(tm_\w+)\s*\(", header, re.M))
        assert len(names) >= 20
        for name in names:
            assert hasattr(lib, name), name
        assert set(_lib.SIGNATURES) >= names
        assert lib.tm_abi_version() == 1

    def test_fft_factor_host_only(self):
        from pytme_amd import _lib

        lib = _lib.lib()
        r = (ctypes.c_int * 16)()
        k = ctypes.c_int()
        assert lib.tm_fft_factor(200, r, 16, ctypes.byref(k)) == 0
        assert list(r)[: k.value] == [8, 5, 5]
        assert int(np.prod(list(r)[: k.value])) == 200


class _CountingAnalyzer:
    """Stand-in analyzer: state (counts array, mapping dict, call number);
    each call adds ``value`` to the counts in place and records it."""

    shareable = True

    def init_state(self):
        from pytme_amd.backends import backend as be

        return be.zeros((4,), dtype=be._float_dtype), {}, 0

    def __call__(self, state, value):
        counts, mapping, n = state
        counts += value
        mapping[n] = value
        return counts, mapping, n + 1

    def result(self, state, **kwargs):
        counts, mapping, n = state
        return counts.clone(), dict(mapping), n

    @classmethod
    def merge(cls, results, **kwargs):
        return len(results)


class TestAnalyzerProxy:
    def test_local_state_follows_the_analyzer(self, cpu_backend):
        from pytme_amd.analyzer.proxy import SharedAnalyzerProxy, StatelessSharedAnalyzerProxy

        proxy = SharedAnalyzerProxy(_CountingAnalyzer, {})
        assert not proxy.is_shared
        proxy(1.0)
        proxy(2.0)
        counts, mapping, n = proxy.result()
        assert n == 2 and mapping == {0: 1.0, 1: 2.0} and float(counts[0]) == 3.0
        assert proxy.state[2] == 2 and isinstance(proxy.analyzer, _CountingAnalyzer)
        assert proxy.merge([1, 2, 3]) == 3
        stateless = StatelessSharedAnalyzerProxy(_CountingAnalyzer, {})
        st = stateless.init_state()
        st = stateless(st, 5.0)
        assert stateless.result(st)[2] == 1

    def test_shared_state_is_published_and_updated_in_place(self, cpu_backend):
        from pytme_amd.analyzer.proxy import SharedAnalyzerProxy

        proxy = SharedAnalyzerProxy(_CountingAnalyzer, {}, shm_handler=object())
        assert proxy.is_shared
        published = proxy.state
        assert isinstance(published[0], tuple)  # the backend's shared-array form
        proxy(1.5)
        proxy(2.5)
        counts, mapping, n = proxy.result()
        # arrays and the manager dict are updated in place; the plain call
        # counter is not part of the shared state, so both calls wrote key 0
        assert float(counts[1]) == 4.0 and mapping == {0: 2.5} and n == 0
        assert proxy.state is published


class TestRotationSets:
    def _write_sets(self, path):
        import yaml
        from scipy.spatial.transform import Rotation

        rng = np.random.default_rng(5)
        sets = {"c48u9.npy": (9, 36.5, 2.9), "c48u27.npy": (27, 20.8, 1.6)}
        for name, (n, angle, cov) in sets.items():
            q = rng.standard_normal((n, 4))
            q /= np.linalg.norm(q, axis=1, keepdims=True)
            np.save(os.path.join(path, name), np.concatenate([q, np.full((n, 1), 1.0 / n)], axis=1))
        with open(os.path.join(path, "metadata.yaml"), "w") as fh:
            yaml.safe_dump({k: list(v) for k, v in sets.items()}, fh)
        return Rotation

    def test_nearest_set_scalar_last(self, tmp_path):
        from pytme_amd.rotations import get_rotation_matrices, load_quaternions_by_angle

        Rotation = self._write_sets(str(tmp_path))
        q, w, n = load_quaternions_by_angle(30.0, data_dir=str(tmp_path))
        assert q.shape == (9, 4) and n == 9 and np.allclose(w, 1 / 9)
        assert load_quaternions_by_angle(15.0, data_dir=str(tmp_path))[0].shape == (27, 4)
        mats = get_rotation_matrices(15.0, data_dir=str(tmp_path))
        want = Rotation.from_quat(np.load(os.path.join(str(tmp_path), "c48u27.npy"))[:, :4]).as_matrix()
        assert np.array_equal(mats, want)
        assert np.allclose(np.einsum("nij,nkj->nik", mats, mats), np.eye(3), atol=1e-12)

    def test_missing_sets_raise(self, tmp_path, monkeypatch):
        from pytme_amd.rotations import get_rotation_matrices

        monkeypatch.delenv("PYTME_ROTATION_SETS", raising=False)
        monkeypatch.setattr("importlib.util.find_spec", lambda name: None)
        with pytest.raises(FileNotFoundError):
            get_rotation_matrices(10.0, data_dir=str(tmp_path))
        monkeypatch.setenv("PYTME_ROTATION_SETS", str(tmp_path))
        self._write_sets(str(tmp_path))
        assert get_rotation_matrices(40.0).shape == (9, 3, 3)

    def test_random_sets(self):
        from pytme_amd.rotations import get_rotation_matrices

        mats = get_rotation_matrices(60.0, dim=3, use_optimized_set=False, seed=1)
        assert mats.shape == (216, 3, 3) and np.array_equal(mats[0], np.eye(3))
        assert np.allclose(np.linalg.det(mats), 1.0)
        two = get_rotation_matrices(30.0, dim=2, seed=2)
        assert two.shape == (12, 2, 2) and np.allclose(np.linalg.det(two), 1.0)


class TestCoordinateMaps:
    @pytest.mark.parametrize("pad", [False, True])
    @pytest.mark.parametrize("tshape", [(7, 6, 5), (8, 8, 8)])
    def test_cart_to_score_indexes_the_raw_volume(self, pad, tshape):
        # the result volume is the raw volume rolled by fourier_shift and
        # cropped by the convolution mode (tme/analyzer/aggregation.py:247-261);
        # cart_to_score(p) must name the raw voxel that lands at p
        from oracle import scan as orc
        from pytme_amd.analyzer._utils import cart_to_score, score_to_cart

        target_shape = (20, 17, 15)
        padded = tuple(n + (t - 1 if pad else 0) for n, t in zip(target_shape, tshape))
        conv, fast, _, shift = orc.fourier_padding(padded, tshape, pad_target=pad)
        raw = np.random.default_rng(3).standard_normal(fast)
        mode = "valid" if pad else "same"
        result = orc.apply_convolution_mode(np.roll(raw, shift, axis=(0, 1, 2)), mode, padded, tshape, conv)
        pos = np.argwhere(np.ones(result.shape, bool))
        kw = dict(fast_shape=fast, targetshape=padded, templateshape=tshape, convolution_shape=conv,
                  fourier_shift=shift, convolution_mode=mode)
        raw_pos, valid = cart_to_score(pos, **kw)
        assert valid.all()
        assert np.array_equal(raw[tuple(raw_pos.T)], result[tuple(pos.T)])
        back, ok = score_to_cart(raw_pos, **kw)
        assert ok.all() and np.array_equal(back, pos)
        # positions outside the result are flagged
        outside = np.array([[-1, 0, 0], [result.shape[0], 0, 0]])
        assert not cart_to_score(outside, **kw)[1].any()


def _stand_in_pytme():
    """A module tree with pyTME's registration surface: a BackendManager
    whose add_backend checks the package's own MatchingBackend base
    (tme/backends/__init__.py:82-166) and a score registry whose FLC /
    correlation entries are functions named flc_scoring / corr_scoring
    (tme/matching_scores.py:1183-1193)."""
    import abc
    import types

    class MatchingBackend(abc.ABC):
        @abc.abstractmethod
        def multiply(self, x1, x2, out=None): ...

        @abc.abstractmethod
        def some_method_the_scan_never_calls(self): ...

    class BackendManager:
        def __init__(self):
            self._BACKEND_REGISTRY, self._backend, self._backend_name = {}, None, None

        def add_backend(self, backend_name, backend_class):
            if not issubclass(backend_class, MatchingBackend):
                raise ValueError("backend_class needs to inherit from MatchingBackend.")
            self._BACKEND_REGISTRY[backend_name] = backend_class

        def change_backend(self, backend_name, **kw):
            self._backend = self._BACKEND_REGISTRY[backend_name](**kw)
            self._backend_name = backend_name

        def available_backends(self):
            return ["numpyfftw"]

    calls = []

    def flc_scoring(**kw):
        calls.append(("flc", kw))
        return "reference loop"

    def corr_scoring(**kw):
        calls.append(("corr", kw))
        return "reference loop"

    def mcc_scoring(**kw):
        return "reference loop"

    tme = types.ModuleType("tme")
    tme.backends = types.ModuleType("tme.backends")
    tme.backends.MatchingBackend = MatchingBackend
    tme.backends.backend = BackendManager()
    tme.matching_scores = types.ModuleType("tme.matching_scores")
    tme.matching_scores.MATCHING_EXHAUSTIVE_REGISTER = {
        "CC": ("cc_setup", corr_scoring), "FLC": ("flc_setup", flc_scoring), "MCC": ("mcc_setup", mcc_scoring),
    }
    return tme, calls


class TestIntegration:
    def test_install_registers_backend_and_routes_scores(self, cpu_backend):
        from pytme_amd.backends.hip_backend import HipBackend
        from pytme_amd.integration import install

        tme, calls = _stand_in_pytme()
        info = install(tme)
        manager = tme.backends.backend
        assert "hip" in manager.available_backends() and "numpyfftw" in manager.available_backends()
        manager.change_backend("hip", device="cpu")
        assert isinstance(manager._backend, HipBackend)
        assert isinstance(manager._backend, tme.backends.MatchingBackend)
        with pytest.raises(NotImplementedError):
            manager._backend.some_method_the_scan_never_calls()
        reg = tme.matching_scores.MATCHING_EXHAUSTIVE_REGISTER
        assert sorted(info["fused_scores"]) == ["CC", "FLC"] and reg["MCC"][1].__name__ == "mcc_scoring"
        assert reg["FLC"][1].__name__ == "flc_scoring"  # wraps keep the name
        # the hip backend on the CPU device declines the fused scan: the
        # reference's own loop runs with the same arguments
        assert reg["FLC"][1](rotations=None, callback=None) == "reference loop"
        assert calls[-1][0] == "flc" and set(calls[-1][1]) == {"rotations", "callback"}
        # a backend whose hook takes the call never reaches the loop
        manager._backend.fused_scoring = lambda score, **kw: ("fused", score)
        n = len(calls)
        assert reg["CC"][1](rotations=None) == ("fused", "corr") and len(calls) == n
        # installing twice wraps nothing twice
        before = dict(reg)
        install(tme)
        assert reg == before and manager.available_backends().count("hip") == 1


class TestPostprocess:
    def test_euler_zyz_matches_scipy(self):
        import warnings

        import torch
        from scipy.spatial.transform import Rotation

        from pytme_amd.postprocess import rotation_matrix_to_euler_zyz
        from pytme_amd.refine import euler_zyz_matrix

        rots = Rotation.random(500, random_state=4)
        got = rotation_matrix_to_euler_zyz(torch.as_tensor(rots.as_matrix())).numpy()
        assert np.abs(np.angle(np.exp(1j * (got - rots.as_euler("zyz"))))).max() < 1e-12
        # gimbal lock: scipy sets the third angle to 0 and the matrix is kept
        locked = Rotation.from_euler("zyz", [[10, 0, 20], [30, 180, 0], [10, 180, 40], [-170, 180, 60]], degrees=True)
        got = rotation_matrix_to_euler_zyz(torch.as_tensor(locked.as_matrix()))
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            want = locked.as_euler("zyz")
        assert np.abs(np.angle(np.exp(1j * (got.numpy() - want)))).max() < 1e-12
        assert np.abs(euler_zyz_matrix(got).numpy() - locked.as_matrix()).max() < 1e-12

    @pytest.mark.parametrize("pad", [False, True])
    def test_device_result_is_the_analyzer_roll_and_crop(self, pad):
        import torch

        from oracle import scan as orc
        from pytme_amd.postprocess import device_result

        target_shape, tshape = (20, 17, 15), (7, 6, 5)
        padded = tuple(n + (t - 1 if pad else 0) for n, t in zip(target_shape, tshape))
        conv, fast, _, shift = orc.fourier_padding(padded, tshape, pad_target=pad)
        raw = np.random.default_rng(2).standard_normal(fast).astype(np.float32)
        rot = np.arange(raw.size, dtype=np.int32).reshape(fast)
        mode = "valid" if pad else "same"
        kw = dict(fourier_shift=shift, convolution_mode=mode, targetshape=padded, templateshape=tshape,
                  convolution_shape=conv)
        s, r = device_result((torch.as_tensor(raw), torch.as_tensor(rot), {}), **kw)
        axes = (0, 1, 2)
        want_s = orc.apply_convolution_mode(np.roll(raw, shift, axes), mode, padded, tshape, conv)
        want_r = orc.apply_convolution_mode(np.roll(rot, shift, axes), mode, padded, tshape, conv)
        assert np.array_equal(s.numpy(), want_s) and np.array_equal(r.numpy(), want_r)


def test_rotation_table_of_both_mapping_kinds():
    import torch

    from pytme_amd.postprocess import rotation_table

    mats = torch.as_tensor(np.random.default_rng(0).standard_normal((5, 3, 3)).astype(np.float32))
    by_index = {i: mats[i] for i in range(5)}
    assert torch.equal(rotation_table(by_index), mats)
    by_bytes = {mats[i].numpy().tobytes(): i for i in (3, 1, 4, 0, 2)}
    assert torch.equal(rotation_table(by_bytes), mats)
    assert rotation_table({}).shape == (0, 3, 3)


def test_fft_radix_plans():
    import ctypes

    from pytme_amd import _lib

    lib = _lib.lib()

    def plan(n, v):
        buf, cnt = (ctypes.c_int * 16)(), ctypes.c_int()
        assert lib.tm_fft_radix_plan(n, v, buf, 16, ctypes.byref(cnt)) == 0
        return [buf[i] for i in range(cnt.value)]

    assert plan(512, 0) == [8, 8, 8] and plan(512, 1) == [4, 4, 4, 8]
    assert plan(100, 0) == [4, 5, 5] and plan(100, 1) == [5, 5, 4]
    assert plan(1024, 0) == [8, 8, 4, 4] and plan(1024, 1) == [4, 4, 8, 8]
    assert plan(256, 1) == [4, 8, 8] and plan(128, 1) == [4, 4, 8] and plan(64, 1) == [4, 4, 4]
    for n in (48, 96, 100, 128, 200, 256, 384, 512, 640, 1024):
        for v in (0, 1):
            assert int(np.prod(plan(n, v))) == n, (n, v)
    assert plan(11, 0) == [] and plan(22, 1) == []
    assert lib.tm_fft_radix_plan(512, 2, (ctypes.c_int * 4)(), 4, ctypes.byref(ctypes.c_int())) != 0
