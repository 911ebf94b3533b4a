"""CPU oracle of the peak callers on the max-score volume: a NumPy/SciPy
restatement of ``PeakCallerMaximumFilter`` / ``PeakCallerSort`` with the
``PeakCaller`` state logic of the reference numpy backend.

TEST INFRASTRUCTURE ONLY.  Nothing in ``pytme_amd`` imports this module; it is
used by ``tests/`` and the ``cpu_baseline`` leg of ``bench.py`` as the
checker, never as the product path.

Functions and the reference lines they restate:
  * ``max_filter_coordinates``  tme/backends/npfftw_backend.py:200-206
  * ``topk_indices``            tme/backends/npfftw_backend.py:176-181
  * ``find_candidate_indices``  tme/external/bindings.cpp:84-124 (int64)
  * ``filter_points_indices``   tme/analyzer/peaks.py:68-106 (exact branch)
  * ``PeakCallerOracle``        tme/analyzer/peaks.py:146-670

Two choices the reference leaves to numpy's unstable sorts are fixed here
(and in the GPU path): equal keys are ordered by lower index (float32 total
order: -0 below +0, NaN above everything), and NaN scores never become
peaks.  On inputs without ties the order is numpy's.

Pinning (tests/test_peaks_oracle.py):
  * ``max_filter_coordinates`` - brute-force box maximum with zero padding on
    small volumes, even and odd edges, negative values and plateaus;
  * ``topk_indices``           - ``numpy.argsort(-v, kind="stable")`` on
    tie-free data, a Python sort by (key, index) with ties, NaN and zeros;
  * ``find_candidate_indices`` - the fixed-point characterisation of greedy
    selection (kept rows pairwise apart, every dropped row near an earlier
    kept row), hand-made distance cases (d2 = 25, 35, 36 at distance 5);
  * the callers                - the reference test's expectations
    (tests/test_analyzer.py:51-73), planted peaks found exactly, ``merge``
    idempotent.
"""

import math
from typing import List, Tuple

import numpy as np
from scipy.ndimage import maximum_filter

__all__ = [
    "max_filter_coordinates",
    "float_keys",
    "topk_indices",
    "find_candidate_indices",
    "filter_points_indices",
    "PeakCallerOracle",
]


def max_filter_coordinates(scores: np.ndarray, min_distance: int) -> np.ndarray:
    """Coordinates (C order) where the score equals its box maximum of edge
    ``min_distance``, zero outside (npfftw_backend.py:200-206)."""
    box = tuple(int(min_distance) for _ in range(scores.ndim))
    filtered = maximum_filter(scores, size=box, mode="constant")
    return np.array(np.nonzero(filtered == scores)).T


def float_keys(values: np.ndarray) -> np.ndarray:
    """uint32 keys with the float32 total order, NaN on top."""
    v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
    u = v.view(np.uint32).astype(np.uint64)
    neg = (u & 0x80000000) != 0
    key = np.where(neg, (~u) & 0xFFFFFFFF, u | 0x80000000)
    key[np.isnan(v)] = 0xFFFFFFFF
    return key.astype(np.int64)


def topk_indices(values: np.ndarray, k: int) -> np.ndarray:
    """Flat indices of the k largest values, largest first, ties by lower
    index (npfftw_backend.py:176-181 with the tie order fixed)."""
    key = float_keys(values)
    order = np.lexsort((np.arange(key.size), -key))
    return order[: int(k)]


def find_candidate_indices(coordinates: np.ndarray, min_distance: int) -> np.ndarray:
    """Greedy selection in input order (bindings.cpp:84-124, the int64
    overload the reference's int64 peak positions select): row i is kept
    unless a kept row j has trunc(sqrt(sum (c_i - c_j)^2)) <= min_distance;
    row 0 is always kept."""
    c = np.asarray(coordinates, dtype=np.int64)
    n = c.shape[0]
    if n == 0:
        return np.zeros(0, dtype=np.int64)
    kept = np.zeros(n, dtype=np.int64)
    rows = np.empty_like(c)
    rows[0], K = c[0], 1
    for i in range(1, n):
        d2 = np.sum((rows[:K] - c[i]) ** 2, axis=1)
        dist = np.trunc(np.sqrt(d2.astype(np.float64))).astype(np.int64)
        if np.any(dist <= min_distance):
            continue
        kept[K], rows[K] = i, c[i]
        K += 1
    return kept[:K]


def filter_points_indices(coordinates: np.ndarray, min_distance: int, scores: np.ndarray = None) -> np.ndarray:
    """tme/analyzer/peaks.py:68-106 (exact branch): sort by score, greedy."""
    coordinates = np.asarray(coordinates)
    n = coordinates.shape[0]
    if min_distance <= 0:
        return np.arange(n)
    order = None
    if scores is not None:
        order = topk_indices(scores, n)
        coordinates = coordinates[order]
    idx = find_candidate_indices(coordinates, min_distance)
    return idx if order is None else order[idx]


class PeakCallerOracle:
    """``kind`` "maximum_filter" (tme/analyzer/peaks.py:662-670) or "sort"
    (:644-659) with the base-class state logic (:146-641)."""

    def __init__(self, kind: str, shape, num_peaks: int = 1000, min_distance: int = 1,
                 min_boundary_distance: int = 0, min_score: float = None, max_score: float = None):
        self.kind = kind
        self.shape = tuple(shape)
        self.num_peaks = int(num_peaks)
        self.min_distance = int(min_distance)
        self.min_boundary_distance = int(min_boundary_distance)
        self.min_score, self.max_score = min_score, max_score

    def init_state(self):
        d = len(self.shape)
        return (np.full((self.num_peaks, d), -1, dtype=np.int64),
                np.repeat(np.eye(d, dtype=np.float32)[None], self.num_peaks, axis=0),
                np.full(self.num_peaks, -1, dtype=np.float32), np.full(self.num_peaks, -1, dtype=np.float32))

    def call_peaks(self, scores):
        if self.kind == "maximum_filter":
            return max_filter_coordinates(scores, self.min_distance)
        k = min(self.num_peaks, scores.size)
        return np.array(np.unravel_index(topk_indices(scores, k), scores.shape)).T

    def _score_window(self, s, valid):
        valid &= s == s
        if self.min_score is not None:
            valid &= s >= self.min_score
        if self.max_score is not None:
            valid &= s <= self.max_score
        return valid

    def get_peak_mask(self, peaks, scores):
        """:241-290 with key = score where valid else 0."""
        if not len(peaks):
            return None
        mbd = self.min_boundary_distance
        valid = np.all((peaks < np.subtract(scores.shape, mbd)) & (peaks >= mbd), axis=1)
        s = scores[tuple(peaks.T)]
        valid = self._score_window(s, valid)
        if valid.sum() == 0:
            return None
        if valid.sum() > self.num_peaks or s.shape[0] > self.num_peaks:
            key = np.where(valid, s, np.float32(0))
            idx = topk_indices(key, min(key.size, 2 * self.num_peaks))
            valid = np.zeros(s.shape, dtype=bool)
            valid[idx] = True
            valid = self._score_window(s, valid)
        return valid

    def update(self, state, positions, rotations, scores, details, offset=None):
        """:531-583."""
        t0, r0, s0, d0 = state
        positions = np.asarray(positions, dtype=np.int64)
        if offset is not None:
            positions = positions + np.asarray(offset, dtype=np.int64)
        pos = np.concatenate((t0, positions))
        rot = np.concatenate((r0, np.asarray(rotations, dtype=np.float32)))
        sc = np.concatenate((s0, np.asarray(scores, dtype=np.float32)))
        de = np.concatenate((d0, np.asarray(details, dtype=np.float32)))
        order = filter_points_indices(pos, self.min_distance, sc)
        top = topk_indices(sc[order], min(order.size, self.num_peaks))
        final = order[top]
        return pos[final], rot[final], sc[final], de[final]

    def __call__(self, state, scores, rotation_matrix=None):
        """:327-384."""
        scores = np.asarray(scores, dtype=np.float32)
        d = scores.ndim
        rotation_matrix = np.eye(d, dtype=np.float32) if rotation_matrix is None else rotation_matrix
        peaks = np.asarray(self.call_peaks(scores), dtype=np.int64)
        valid = self.get_peak_mask(peaks, scores)
        if valid is None:
            return state
        peaks = peaks[valid]
        s = scores[tuple(peaks.T)]
        rot = np.repeat(np.asarray(rotation_matrix, dtype=np.float32)[None], peaks.shape[0], axis=0)
        return self.update(state, peaks, rot, s, np.full(s.shape, -1, dtype=np.float32))

    @staticmethod
    def result(state):
        """:585-641 without coordinate maps: drop rows with negative positions."""
        pos, rot, sc, de = state
        valid = np.all(pos >= 0, axis=1)
        return pos[valid], rot[valid], sc[valid], de[valid]

    def merge(self, results: List[Tuple]):
        """:387-426."""
        state = self.init_state()
        for pos, rot, sc, de in results:
            state = self.update(state, pos, rot, sc, de)
        return state


def brute_force_max_filter(scores: np.ndarray, size: int) -> np.ndarray:
    """Box maximum of edge ``size`` (window i - size//2 .. i + size - 1 -
    size//2) with zero outside, by explicit loops; for pinning only."""
    lo = size // 2
    hi = size - 1 - lo
    out = np.empty_like(scores)
    n0, n1, n2 = scores.shape
    for z in range(n0):
        for y in range(n1):
            for x in range(n2):
                m = -math.inf
                for dz in range(-lo, hi + 1):
                    for dy in range(-lo, hi + 1):
                        for dx in range(-lo, hi + 1):
                            zz, yy, xx = z + dz, y + dy, x + dx
                            if 0 <= zz < n0 and 0 <= yy < n1 and 0 <= xx < n2:
                                v = scores[zz, yy, xx]
                            else:
                                v = 0.0
                            m = max(m, v)
                out[z, y, x] = m
    return out
