"""Peak extraction on the GPU vs the oracle (oracle/peaks.py): the box
maximum filter, top-k selection and the greedy distance filter through the
C-ABI, and PeakCallerMaximumFilter / PeakCallerSort end to end, bit-exact
(integer coordinates, identical float32 scores, identical order)."""

import numpy as np
import pytest

from oracle import peaks as op

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    from pytme_amd.backends import backend

    backend.change_backend("hip", device="cuda")
    return backend


def _volume(rng, shape, kind):
    if kind == "normal":
        return rng.normal(size=shape).astype(np.float32)
    if kind == "plateau":
        return np.round(2 * rng.normal(size=shape)).astype(np.float32)
    return rng.random(shape).astype(np.float32)


@pytest.mark.parametrize(
    "shape", [(37, 41, 53), (64, 64, 64), (5, 300, 7), (1, 1, 1), (3, 1, 200), (2, 70, 130), (70, 33, 200)]
)
@pytest.mark.parametrize("size", [1, 2, 3, 5, 8, 9, 12, 16, 17, 24, 32, 33, 64])
@pytest.mark.parametrize("kind", ["normal", "plateau"])
def test_max_filter_coordinates(be, shape, size, kind):
    import torch

    rng = np.random.default_rng(sum(shape) * 1000 + size * 10 + (kind == "plateau"))
    s = _volume(rng, shape, kind)
    print(f"max filter {shape} size {size} {kind}", flush=True)
    got = be.max_filter_coordinates(torch.as_tensor(s).cuda(), size).cpu().numpy()
    want = op.max_filter_coordinates(s, size)
    np.testing.assert_array_equal(got, want)


def _values(rng, n, kind):
    if kind == "distinct":
        return ((rng.permutation(n) - n // 2) * 0.5).astype(np.float32)
    v = rng.integers(-4, 5, size=n).astype(np.float32)
    m = max(1, n // 50)
    v[rng.integers(0, n, m)] = np.nan
    v[rng.integers(0, n, m)] = -0.0
    v[rng.integers(0, n, m)] = np.inf
    v[rng.integers(0, n, m)] = -np.inf
    return v


@pytest.mark.parametrize("n", [1, 100, 4096, 4097, 100003, 2000000])
@pytest.mark.parametrize("kind", ["distinct", "ties"])
def test_topk_indices(be, n, kind):
    import torch

    rng = np.random.default_rng(n + (kind == "ties"))
    v = _values(rng, n, kind)
    vd = torch.as_tensor(v).cuda()
    want = op.topk_indices(v, n)
    for k in sorted({1, 17, min(n, 2000), min(n, 8192), min(n, 8193), min(n, 12288), min(n, 65536)}):
        print(f"topk n {n} k {k} {kind}", flush=True)
        got = be.topk_indices(vd, k)[0].cpu().numpy()
        np.testing.assert_array_equal(got, want[:k])


def test_topk_unravels_like_numpy(be):
    import torch

    rng = np.random.default_rng(3)
    v = rng.random((20, 30, 40)).astype(np.float32)
    print("topk unravel (20,30,40)", flush=True)
    got = be.topk_indices(torch.as_tensor(v).cuda(), 50)
    want = np.unravel_index(op.topk_indices(v, 50), v.shape)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g.cpu().numpy(), w)


@pytest.mark.parametrize("n", [1, 2, 50, 5000, 12288, 16384, 20000])
@pytest.mark.parametrize("md", [1, 5, 12])
@pytest.mark.parametrize("layout", ["spread", "clustered"])
def test_greedy_distance_filter(be, n, md, layout):
    """Parallel rounds (n <= 16384) and the sequential pass (larger n); the
    clustered layout gives candidates with more than 8 earlier neighbours."""
    import torch

    rng = np.random.default_rng(n * 31 + md + (layout == "clustered"))
    if layout == "spread":
        c = rng.integers(-1, 120, size=(n, 3)).astype(np.int64)
    else:
        centres = rng.integers(0, 200, size=(max(1, n // 200), 3))
        c = (centres[rng.integers(0, centres.shape[0], n)] + rng.integers(-4, 5, size=(n, 3))).astype(np.int64)
    want = op.find_candidate_indices(c, md)
    cd = torch.as_tensor(c).cuda()
    for max_keep in (1, 100, 4096, None):
        if max_keep is None and n > 16384:
            continue
        print(f"distance filter n {n} md {md} {layout} max_keep {max_keep}", flush=True)
        got = be.greedy_distance_filter(cd, md, max_keep).cpu().numpy()
        np.testing.assert_array_equal(got, want if max_keep is None else want[:max_keep])


def _compare_state(got, want):
    for g, w in zip(got, want):
        np.testing.assert_array_equal(np.asarray(g), np.asarray(w))


CALLER_CASES = [
    # (num_peaks, min_distance, min_boundary_distance, min_score, max_score)
    (1, 5, 0, None, None),
    (100, 5, 0, None, None),
    (100, 5, 0, 0.5, None),
    (1000, 5, 4, None, None),
    (300, 3, 2, 0.2, 0.95),
    (4096, 2, 0, None, None),
    (50, 0, 0, None, None),
]


@pytest.fixture(params=["device_fold", "primitives"])
def step_path(request, be, monkeypatch):
    """Each caller step through tm_peak_caller_fold, or (fold disabled) the
    primitive-by-primitive flow."""
    if request.param == "primitives":
        monkeypatch.setattr(be._active(), "fold_max_rows", 0)
    return request.param


@pytest.mark.parametrize("kind", ["maximum_filter", "sort"])
@pytest.mark.parametrize("case", CALLER_CASES)
def test_peak_callers_match_oracle(be, step_path, kind, case):
    """tests/test_analyzer.py:51-73's setup (100^3 uniform, min_distance 5)
    and variations; also the merge of the result."""
    import torch

    from pytme_amd.analyzer import PeakCallerMaximumFilter, PeakCallerSort

    num_peaks, md, mbd, lo, hi = case
    rng = np.random.default_rng(123)
    data = rng.random((100, 100, 100)).astype(np.float32)
    cls = PeakCallerMaximumFilter if kind == "maximum_filter" else PeakCallerSort
    kw = dict(shape=data.shape, num_peaks=num_peaks, min_distance=md, min_boundary_distance=mbd, min_score=lo,
              max_score=hi)
    print(f"peak caller {kind} {case} {step_path}", flush=True)
    pc = cls(**kw)
    state = pc(pc.init_state(), torch.as_tensor(data).cuda(), rotation_matrix=np.eye(3))
    res = pc.result(state)
    ref = op.PeakCallerOracle(kind, **kw)
    want = ref.result(ref(ref.init_state(), data))
    _compare_state(res, want)
    # the sort caller's top voxels can all lie above max_score: then no peaks
    assert res[0].shape[0] <= num_peaks and (hi is not None or res[0].shape[0] >= 1)
    if lo is not None:
        assert np.all(res[2] >= lo)
    merged = cls.merge(results=[res], **kw)
    _compare_state(merged, ref.merge([want]))


EDGE_CASES = [
    # (volume, num_peaks, min_distance, min_boundary_distance, min_score, max_score)
    ("uniform", 100, 5, 0, 2.0, None),     # no valid candidate: the state stays as it is
    ("negative", 50, 3, 10, None, None),   # preselection keys 0 > negative scores: boundary peaks come back
    ("nan", 200, 4, 2, None, None),        # NaN scores never become peaks
    ("sparse", 500, 6, 0, None, None),     # fewer candidates than num_peaks: no preselection
    ("uniform", 2000, 1, 3, 0.3, 0.9),     # 2 num_peaks + state rows close to the device limit
]


@pytest.mark.parametrize("kind", ["maximum_filter", "sort"])
@pytest.mark.parametrize("case", EDGE_CASES)
def test_peak_caller_edge_cases(be, step_path, kind, case):
    import torch

    from pytme_amd.analyzer import PeakCallerMaximumFilter, PeakCallerSort

    vol, num_peaks, md, mbd, lo, hi = case
    rng = np.random.default_rng(77)
    data = rng.random((60, 50, 40)).astype(np.float32)
    if vol == "negative":
        data = data - 1.5
    elif vol == "nan":
        data[rng.random(data.shape) < 0.01] = np.nan
    elif vol == "sparse":
        data = np.zeros((60, 50, 40), dtype=np.float32)
        data[tuple(rng.integers(0, 40, size=(3, 60)))] = rng.random(60).astype(np.float32) + 0.5
    cls = PeakCallerMaximumFilter if kind == "maximum_filter" else PeakCallerSort
    kw = dict(shape=data.shape, num_peaks=num_peaks, min_distance=md, min_boundary_distance=mbd, min_score=lo,
              max_score=hi)
    print(f"peak caller edge case {kind} {case} {step_path}", flush=True)
    pc = cls(**kw)
    state = pc(pc.init_state(), torch.as_tensor(data).cuda(), rotation_matrix=np.eye(3))
    ref = op.PeakCallerOracle(kind, **kw)
    want_state = ref(ref.init_state(), data)
    # the raw state too (placeholder rows included), not only the result
    for g, w in zip(state, want_state):
        np.testing.assert_array_equal(np.asarray(g.cpu() if hasattr(g, "cpu") else g), np.asarray(w))
    _compare_state(pc.result(state), ref.result(want_state))


def test_two_volumes_into_one_state(be, step_path):
    """Two score volumes (two rotations) folded into one state, then result:
    the state logic across calls."""
    import torch

    from pytme_amd.analyzer import PeakCallerMaximumFilter

    rng = np.random.default_rng(9)
    a = rng.random((60, 61, 62)).astype(np.float32)
    b = rng.random((60, 61, 62)).astype(np.float32)
    ra = np.eye(3, dtype=np.float32)
    rb = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], dtype=np.float32)
    kw = dict(shape=a.shape, num_peaks=200, min_distance=4)
    print(f"two volumes into one state ({step_path})", flush=True)
    pc = PeakCallerMaximumFilter(**kw)
    st = pc(pc.init_state(), torch.as_tensor(a).cuda(), rotation_matrix=ra)
    st = pc(st, torch.as_tensor(b).cuda(), rotation_matrix=rb)
    ref = op.PeakCallerOracle("maximum_filter", **kw)
    rst = ref(ref.init_state(), a, ra)
    rst = ref(rst, b, rb)
    _compare_state(pc.result(st), ref.result(rst))


def test_peaks_of_the_scan_result(be):
    """The caller flow of scripts/postprocess.py:482-533: scan_subsets of the
    512x512x200 / 48^3 FLC workload (64 rotations, fused path) -> result
    scores -> PeakCallerMaximumFilter(num_peaks 1000, min_distance 5),
    identical to the oracle on the same score volume."""
    import torch

    from pytme_amd.analyzer import MaxScoreOverRotations, PeakCallerMaximumFilter
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, scan_subsets
    from pytme_amd.synthetic import make_workload

    target, template, mask, rotations, filters, spec = make_workload("flc_512x512x200_t48", n_rotations=64)
    print("scan 64 rotations of 512x512x200 / 48^3 -> peak calling", flush=True)
    md = MatchingData(target, template, template_mask=mask, rotations=rotations)
    md.template_filter, md.target_filter = filters
    setup, scoring = MATCHING_EXHAUSTIVE_REGISTER[spec["score"]]
    scores, _, rot_array, mapping = scan_subsets(
        matching_data=md, matching_setup=setup, matching_score=scoring, callback_class=MaxScoreOverRotations,
        callback_class_args={"score_threshold": 0.0}, job_schedule=(1, 1), interpolation_order=1,
        pad_template_filter=False,
    )
    scores = np.ascontiguousarray(scores, dtype=np.float32)
    kw = dict(shape=scores.shape, num_peaks=1000, min_distance=5, min_boundary_distance=24)
    pc = PeakCallerMaximumFilter(**kw)
    res = pc.result(pc(pc.init_state(), torch.as_tensor(scores).cuda(), rotation_matrix=np.eye(3)))
    ref = op.PeakCallerOracle("maximum_filter", **kw)
    want = ref.result(ref(ref.init_state(), scores))
    _compare_state(res, want)
    assert res[0].shape[0] == 1000
    # postprocess.py:543-566: rotation of each peak from the rotation array
    # and mapping, as zyz Euler angles, into Orientations; STAR and the
    # result bundle round-trip
    import os
    import tempfile

    from pytme_amd import Orientations
    from pytme_amd.matching_utils import load_pickle, write_pickle
    from pytme_amd.rotations import euler_from_rotationmatrix

    idx = rot_array[tuple(res[0].T)]
    assert np.all(idx >= 0)
    angles = np.vstack([euler_from_rotationmatrix(np.asarray(mapping[int(i)])) for i in idx])
    orientations = Orientations(translations=res[0], rotations=angles, scores=res[2], details=res[3])
    with tempfile.TemporaryDirectory() as d:
        star = os.path.join(d, "peaks.star")
        orientations.to_file(star)
        back = Orientations.from_file(star)
        np.testing.assert_array_equal(back.translations, res[0].astype(np.float32))
        np.testing.assert_allclose(back.rotations, angles, atol=1e-4)
        bundle = os.path.join(d, "result.pickle")
        write_pickle((scores, np.zeros(3, int), rot_array, mapping), bundle)
        s2, _, r2, m2 = load_pickle(bundle)
        np.testing.assert_array_equal(s2, scores)
        np.testing.assert_array_equal(r2, rot_array)
        assert set(m2) == set(mapping)
