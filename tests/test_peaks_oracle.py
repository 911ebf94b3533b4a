"""Pins the peak-caller oracle (oracle/peaks.py) against independent checks,
and the host-side argument handling of the product peak callers."""

import math

import numpy as np
import pytest

from oracle import peaks as op


@pytest.mark.parametrize("size", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("plateau", [False, True])
def test_max_filter_matches_brute_force(size, plateau):
    """scipy maximum_filter(mode="constant") == explicit box maximum with zero
    outside (window i - size//2 .. i + size - 1 - size//2), negative values
    included; plateaus make every tied voxel a peak."""
    rng = np.random.default_rng(size + 10 * plateau)
    s = rng.normal(size=(6, 7, 9)).astype(np.float32)
    if plateau:
        s = np.round(s).astype(np.float32)
    want = np.array(np.nonzero(op.brute_force_max_filter(s, size) == s)).T
    got = op.max_filter_coordinates(s, size)
    np.testing.assert_array_equal(got, want)
    if size == 1:
        assert got.shape[0] == s.size
    # a negative voxel next to the border is never a peak once the window
    # reaches outside (the outside counts as 0)
    if size >= 2:
        for z, y, x in got:
            if s[z, y, x] < 0:
                lo, hi = size // 2, size - 1 - size // 2
                inside = (z - lo >= 0 and z + hi < 6 and y - lo >= 0 and y + hi < 7 and x - lo >= 0 and x + hi < 9)
                assert inside


def test_topk_matches_argsort_without_ties():
    rng = np.random.default_rng(0)
    v = ((rng.permutation(5000) - 2500) * 0.25).astype(np.float32)
    assert np.unique(v).size == v.size
    for k in (1, 10, 777, 5000):
        np.testing.assert_array_equal(op.topk_indices(v, k), np.argsort(-v, kind="stable")[:k])
        # the same set numpy's argpartition selects (npfftw_backend.py:176-181)
        part = np.argpartition(v, -k)[-k:]
        np.testing.assert_array_equal(np.sort(op.topk_indices(v, k)), np.sort(part))


def test_topk_tie_order_nan_and_zeros():
    rng = np.random.default_rng(1)
    v = rng.integers(-3, 4, size=400).astype(np.float32)
    v[rng.integers(0, 400, 7)] = np.nan
    v[rng.integers(0, 400, 9)] = -0.0
    v[rng.integers(0, 400, 5)] = np.inf
    v[rng.integers(0, 400, 5)] = -np.inf

    def key(i):
        x = float(v[i])
        if math.isnan(x):
            return (0, 0.0, 0, i)
        return (1, -x, 1 if (x == 0 and math.copysign(1, x) < 0) else 0, i)

    want = sorted(range(v.size), key=key)
    np.testing.assert_array_equal(op.topk_indices(v, v.size), want)
    np.testing.assert_array_equal(op.topk_indices(v, 13), want[:13])


def test_find_candidate_indices_distance_rule():
    """int64 trunc(sqrt(d2)) <= min_distance drops a point: d2 = 25, 33 and
    35 are dropped at min_distance 5, d2 = 36 is kept."""
    base = np.array([[0, 0, 0]])
    for other, kept in (([5, 0, 0], False), ([5, 3, 1], False), ([4, 4, 1], False), ([6, 0, 0], True),
                        ([4, 4, 2], True)):
        c = np.vstack([base, other])
        assert (op.find_candidate_indices(c, 5).size == 2) == kept, other
    # the closed form the GPU filter uses
    for md in range(0, 40):
        for d2 in range(0, 2000):
            assert (int(math.sqrt(d2)) <= md) == (d2 < (md + 1) ** 2)


@pytest.mark.parametrize("md", [1, 3, 5, 10])
def test_find_candidate_indices_fixed_point(md):
    """Greedy in order is the unique set with: row 0 kept, kept rows pairwise
    farther than md (integer rule), every dropped row within md of an earlier
    kept row."""
    rng = np.random.default_rng(md)
    c = rng.integers(0, 40, size=(400, 3))
    kept = op.find_candidate_indices(c, md)
    assert kept[0] == 0
    ks = set(kept.tolist())

    def near(i, j):
        return int(math.sqrt(int(np.sum((c[i] - c[j]) ** 2)))) <= md

    for a in kept:
        for b in kept:
            if a < b:
                assert not near(a, b)
    for i in range(c.shape[0]):
        if i not in ks:
            assert any(near(i, j) for j in kept if j < i)


@pytest.mark.parametrize("kind", ["maximum_filter", "sort"])
@pytest.mark.parametrize("num_peaks", [1, 100])
@pytest.mark.parametrize("min_score", [None, 0.5])
def test_reference_caller_expectations(kind, num_peaks, min_score):
    """tests/test_analyzer.py:51-73 on 100^3 uniform data, min_distance 5."""
    rng = np.random.default_rng(123)
    data = rng.random((100, 100, 100)).astype(np.float32)
    pc = op.PeakCallerOracle(kind, data.shape, num_peaks=num_peaks, min_distance=5, min_score=min_score)
    state = pc(pc.init_state(), data)
    pos, rot, sc, de = pc.result(state)
    assert 1 <= pos.shape[0] <= num_peaks
    np.testing.assert_array_equal(data[tuple(pos.T)], sc)
    if min_score is not None:
        assert np.all(sc >= min_score)
    assert np.all(np.diff(sc) <= 0)
    for i in range(pos.shape[0]):
        for j in range(i):
            assert int(math.sqrt(int(np.sum((pos[i] - pos[j]) ** 2)))) > 5
    if kind == "maximum_filter":
        mf = op.maximum_filter(data, size=5, mode="constant")
        assert np.all(mf[tuple(pos.T)] == sc)
    else:
        # the sort caller's first peak is the global maximum
        assert sc[0] == data.max()


def test_planted_peaks_found_exactly():
    rng = np.random.default_rng(4)
    shape = (64, 64, 64)
    data = (0.01 * rng.normal(size=shape)).astype(np.float32)
    zz, yy, xx = np.indices(shape)
    centres = [(10, 12, 14), (50, 20, 33), (30, 45, 9), (20, 30, 50), (45, 50, 45), (8, 52, 30)]
    amps = [1.0, 0.9, 0.8, 0.7, 0.6, 0.5]
    for (z, y, x), a in zip(centres, amps):
        data += (a * np.exp(-((zz - z) ** 2 + (yy - y) ** 2 + (xx - x) ** 2) / 4.0)).astype(np.float32)
    pc = op.PeakCallerOracle("maximum_filter", shape, num_peaks=6, min_distance=5)
    pos, _, sc, _ = pc.result(pc(pc.init_state(), data))
    np.testing.assert_array_equal(pos, np.asarray(centres))


def test_merge_is_idempotent():
    rng = np.random.default_rng(5)
    data = rng.random((48, 50, 52)).astype(np.float32)
    pc = op.PeakCallerOracle("maximum_filter", data.shape, num_peaks=50, min_distance=4)
    res = pc.result(pc(pc.init_state(), data))
    merged = pc.result(pc.merge([res]))
    for a, b in zip(res, merged):
        np.testing.assert_array_equal(a, b)


# ----------------------------------------------------------- product, host side


def test_peak_caller_arguments():
    """tests/test_analyzer.py:35-49: abstract base, parameter checks."""
    from pytme_amd.analyzer import PeakCaller, PeakCallerMaximumFilter, PeakCallerSort

    with pytest.raises(TypeError):
        PeakCaller(shape=(10, 10, 10), num_peaks=100, min_distance=5)
    for cls in (PeakCallerSort, PeakCallerMaximumFilter):
        cls(shape=(10, 10, 10), num_peaks=100, min_distance=5)
        with pytest.raises(ValueError):
            cls(shape=(10, 10, 10), num_peaks=0, min_distance=5)
        with pytest.raises(ValueError):
            cls(shape=(10, 10, 10), num_peaks=-1, min_distance=5)
        with pytest.raises(ValueError):
            cls(shape=(10, 10, 10), num_peaks=10, min_distance=-1)
        with pytest.raises(ValueError):
            cls(shape=(10, 10, 10), num_peaks=10, min_boundary_distance=-1)


def test_peak_primitives_need_the_gpu():
    """No CPU path: on a CPU device the peak primitives raise."""
    from pytme_amd.backends import backend as be

    be.change_backend("hip", device="cpu")
    try:
        with pytest.raises(RuntimeError):
            be.max_filter_coordinates(np.zeros((4, 4, 4), np.float32), 3)
        with pytest.raises(RuntimeError):
            be.topk_indices(np.zeros(10, np.float32), 3)
        with pytest.raises(RuntimeError):
            be.greedy_distance_filter(np.zeros((3, 3), np.int64), 3)
    finally:
        be.change_backend("hip")
