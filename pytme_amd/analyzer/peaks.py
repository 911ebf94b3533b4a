"""Peak callers on the max-score volume.

Re-creates ``PeakCaller``, ``PeakCallerSort`` and ``PeakCallerMaximumFilter``
(tme/analyzer/peaks.py:146-670) with ``filter_points_indices`` /
``filter_points`` (:68-114).  The control flow (state of ``num_peaks``
placeholder rows at -1, the boundary / score mask, the top-``2 num_peaks``
preselection, the greedy minimum-distance filter in score order, the final
top-``num_peaks``, ``result`` and ``merge``) is the reference's; the volume
and selection work runs on the GPU through the backend:

* ``max_filter_coordinates`` -- box maximum filter + peak test + ordered
  compaction (``tm_peaks_max_filter``, ``tm_peaks_write``);
* ``topk_indices`` -- radix select + sort (``tm_topk_f32``);
* ``greedy_distance_filter`` -- the distance filter (``tm_peak_filter``),
  with the reference's integer rule ``trunc(sqrt(d2)) <= min_distance``;
* ``fold_peaks`` -- the rest of one ``__call__`` (mask, preselection,
  merge, sort, distance filter) in one device call
  (``tm_peak_caller_fold``) when the state and the candidates fit it; the
  step below is the same computation in backend primitives.

Deterministic ties: where the reference's result depends on numpy's
unstable ``argpartition`` / ``argsort`` order (equal scores), the lower index
wins here.  NaN scores never become peaks (the reference re-admits them
through ``peak_scores * valid`` when no ``min_score`` is given).  Stacks of
score volumes (``batch_dims``) and the bucket filter the reference uses for
more than 1e5 points are not supported.
"""

from abc import abstractmethod
from typing import List, Tuple

import numpy as np

from ..backends import backend as be
from ._utils import score_to_cart
from .base import AbstractAnalyzer

__all__ = [
    "PeakCaller",
    "PeakCallerSort",
    "PeakCallerMaximumFilter",
    "filter_points",
    "filter_points_indices",
]


def _torch():
    import torch

    return torch


def filter_points_indices(coordinates, min_distance: float, bucket_cutoff: int = 1e5, batch_dims=None,
                          scores=None, max_keep: int = None):
    """Indices of the rows of ``coordinates`` kept by the greedy
    minimum-distance rule, highest score first when ``scores`` are given
    (tme/analyzer/peaks.py:68-106, the exact branch).  ``max_keep`` stops the
    filter after that many kept rows (the callers only use the first
    ``num_peaks``)."""
    if batch_dims is not None:
        raise NotImplementedError("batch_dims is not supported by the hip peak callers")
    coordinates = be.to_backend_array(coordinates)
    n = int(coordinates.shape[0])
    if min_distance <= 0:
        return be.arange(n)
    if n == 0:
        return ()
    order = None
    if scores is not None:
        order = be.topk_indices(be.to_backend_array(scores).reshape(-1), n)[0]
        coordinates = coordinates[order]
    keep = be.greedy_distance_filter(coordinates, int(min_distance), max_keep)
    return keep if order is None else order[keep]


def filter_points(coordinates, min_distance: Tuple[int], batch_dims: Tuple[int] = None):
    """Rows of ``coordinates`` kept by :func:`filter_points_indices` (:109-114)."""
    return coordinates[filter_points_indices(coordinates, min_distance, batch_dims=batch_dims)]


class PeakCaller(AbstractAnalyzer):
    """Base class of the peak callers (tme/analyzer/peaks.py:146-641).

    Parameters are the reference's: ``shape`` of the score space,
    ``num_peaks`` (> 0), ``min_distance`` (>= 0) between peaks,
    ``min_boundary_distance`` (>= 0) to the volume faces, optional
    ``min_score`` / ``max_score``.  The state is (positions (n, d) int64,
    rotations (n, d, d), scores (n,), details (n,)) on the device.
    """

    def __init__(self, shape: Tuple[int], num_peaks: int = 1000, min_distance: int = 1,
                 min_boundary_distance: int = 0, min_score: float = None, max_score: float = None,
                 batch_dims: Tuple[int] = None, shm_handler: object = None, **kwargs):
        if num_peaks <= 0:
            raise ValueError("num_peaks has to be larger than 0.")
        if min_distance < 0:
            raise ValueError("min_distance has to be non-negative.")
        if min_boundary_distance < 0:
            raise ValueError("min_boundary_distance has to be non-negative.")
        if batch_dims is not None:
            raise NotImplementedError("batch_dims is not supported by the hip peak callers")
        self.shape = shape
        self.num_peaks = int(num_peaks)
        self.min_distance = int(min_distance)
        self.min_boundary_distance = int(min_boundary_distance)
        self.batch_dims = None
        self.min_score, self.max_score = min_score, max_score

    @abstractmethod
    def call_peaks(self, scores, **kwargs):
        """Peak positions (n, d) and details (n,) or None in ``scores``."""

    def init_state(self):
        t = _torch()
        ndim = len(self.shape)
        translations = be.full((self.num_peaks, ndim), fill_value=-1, dtype=t.int64)
        rotations = be.zeros((self.num_peaks, ndim, ndim), dtype=be._float_dtype)
        diag = t.arange(ndim, device=rotations.device)
        rotations[:, diag, diag] = 1.0
        scores = be.full((self.num_peaks,), fill_value=-1, dtype=be._float_dtype)
        details = be.full((self.num_peaks,), fill_value=-1, dtype=be._float_dtype)
        return translations, rotations, scores, details

    def _get_peak_mask(self, peaks, scores):
        """Boundary, NaN and score-window mask of the candidate peaks, with the
        top-``2 num_peaks`` preselection (tme/analyzer/peaks.py:241-290)."""
        t = _torch()
        if not len(peaks):
            return None
        dims = t.as_tensor(tuple(scores.shape), dtype=t.int64, device=peaks.device)
        mbd = self.min_boundary_distance
        valid = ((peaks < dims - mbd) & (peaks >= mbd)).all(dim=1)
        peak_scores = scores[tuple(peaks.T)]
        valid &= peak_scores == peak_scores
        if self.min_score is not None:
            valid &= peak_scores >= self.min_score
        if self.max_score is not None:
            valid &= peak_scores <= self.max_score
        n_valid = int(valid.sum().item())
        if n_valid == 0:
            return None
        if n_valid > self.num_peaks or peak_scores.shape[0] > self.num_peaks:
            key = t.where(valid, peak_scores, t.zeros((), dtype=peak_scores.dtype, device=peak_scores.device))
            idx = self._top_peaks(peaks, scores=key, num_peaks=2 * self.num_peaks)
            valid = t.zeros(peak_scores.shape, dtype=t.bool, device=peaks.device)
            valid[idx] = True
            valid &= peak_scores == peak_scores
            if self.min_score is not None:
                valid &= peak_scores >= self.min_score
            if self.max_score is not None:
                valid &= peak_scores <= self.max_score
        return valid

    def __call__(self, state: Tuple, scores, rotation_matrix, **kwargs) -> Tuple:
        """Fold the peaks of ``scores`` into ``state`` (:327-384)."""
        t = _torch()
        scores = be.to_backend_array(scores)
        if scores.dtype != be._float_dtype or not scores.is_contiguous():
            scores = scores.to(be._float_dtype).contiguous()
        rotation_matrix = be.to_backend_array(rotation_matrix).to(be._float_dtype)
        peak_positions, peak_details = self.call_peaks(
            scores=scores, rotation_matrix=rotation_matrix, min_score=self.min_score, max_score=self.max_score,
            **kwargs,
        )
        if peak_positions is None:
            return state
        peak_positions = be.to_backend_array(peak_positions).to(t.int64)
        if peak_details is None and len(peak_positions):
            # the whole step in one device call when the backend has it
            fold = getattr(be, "fold_peaks", None)
            new_state = None if fold is None else fold(
                state, peak_positions, scores, rotation_matrix, self.num_peaks, self.min_distance,
                self.min_boundary_distance, self.min_score, self.max_score,
            )
            if new_state is not None:
                return new_state
        valid = self._get_peak_mask(peaks=peak_positions, scores=scores)
        if valid is None:
            return state
        peak_positions = peak_positions[valid]
        peak_scores = scores[tuple(peak_positions.T)]
        if peak_details is not None:
            peak_details = be.to_backend_array(peak_details)[valid].to(be._float_dtype)
        else:
            peak_details = be.full(peak_scores.shape, fill_value=-1, dtype=be._float_dtype)
        rotations = rotation_matrix.reshape(1, *rotation_matrix.shape).repeat(peak_positions.shape[0], 1, 1)
        return self._update(state, peak_positions=peak_positions, peak_details=peak_details,
                            peak_scores=peak_scores, peak_rotations=rotations)

    @classmethod
    def merge(cls, results: List[Tuple] = (), **kwargs) -> Tuple:
        """Combine ``result`` tuples of several callers (:387-426); returns
        host arrays (positions, rotations, scores, details)."""
        if "shape" not in kwargs:
            kwargs["shape"] = tuple(1 for _ in range(np.asarray(results[0][0]).shape[1]))
        base = cls(**kwargs)
        state = base.init_state()
        for result in results:
            if len(result) == 0:
                continue
            positions, rotations, scores, details = result
            state = base._update(
                state,
                peak_positions=be.to_backend_array(np.asarray(positions)),
                peak_details=be.to_backend_array(np.asarray(details)),
                peak_scores=be.to_backend_array(np.asarray(scores)),
                peak_rotations=be.to_backend_array(np.asarray(rotations)),
                offset=kwargs.get("offset", None),
            )
        return tuple(be.to_numpy_array(x) for x in state)

    def _top_peaks(self, positions, scores, num_peaks: int = None):
        """Indices of the ``num_peaks`` highest scores, highest first (:501-529)."""
        num_peaks = be.size(scores) if not num_peaks else num_peaks
        top_n = min(be.size(scores), num_peaks)
        return be.topk_indices(scores, top_n)[0]

    def _update(self, state, peak_positions, peak_details, peak_scores, peak_rotations, offset=None):
        """Concatenate, distance-filter in score order and keep the top
        ``num_peaks`` (:531-583)."""
        t = _torch()
        translations, rotations, scores, details = (be.to_backend_array(x) for x in state)
        peak_positions = be.to_backend_array(peak_positions).to(t.int64)
        if offset is not None:
            peak_positions = peak_positions + be.to_backend_array(np.asarray(offset)).to(t.int64)
        positions = t.cat((translations.to(t.int64), peak_positions))
        rotations = t.cat((rotations.to(be._float_dtype), be.to_backend_array(peak_rotations).to(be._float_dtype)))
        scores = t.cat((scores.to(be._float_dtype), be.to_backend_array(peak_scores).to(be._float_dtype)))
        details = t.cat((details.to(be._float_dtype), be.to_backend_array(peak_details).to(be._float_dtype)))
        if self.min_distance > 0 and translations.shape[0] > 1:
            # placeholder rows (position -1, score -1) are identical: each one
            # after the first lies at distance 0 from it and shares its fate
            # or is dropped, so one of them decides the filter's result
            n0 = translations.shape[0]
            filler = (positions[:n0] == -1).all(dim=1) & (scores[:n0] == -1)
            drop = filler & (t.cumsum(filler.to(t.int64), 0) > 1)
            if bool(drop.any()):
                keep = t.cat((~drop, t.ones(positions.shape[0] - n0, dtype=t.bool, device=positions.device)))
                positions, rotations, scores, details = positions[keep], rotations[keep], scores[keep], details[keep]
        distance_order = filter_points_indices(positions, self.min_distance, batch_dims=None, scores=scores,
                                               max_keep=self.num_peaks)
        top = self._top_peaks(positions[distance_order], scores[distance_order], self.num_peaks)
        final = distance_order[top]
        return positions[final], rotations[final], scores[final], details[final]

    def result(self, state, fast_shape: Tuple[int] = None, targetshape: Tuple[int] = None,
               templateshape: Tuple[int] = None, convolution_shape: Tuple[int] = None,
               fourier_shift: Tuple[int] = None, convolution_mode: str = None, **kwargs):
        """Map positions to target coordinates and drop the placeholder and
        out-of-range rows (:585-641); host arrays."""
        translations, rotations, scores, details = (be.to_numpy_array(x) for x in state)
        positions, valid = score_to_cart(
            positions=translations, fast_shape=fast_shape, targetshape=targetshape, templateshape=templateshape,
            convolution_shape=convolution_shape, fourier_shift=fourier_shift, convolution_mode=convolution_mode,
            **kwargs,
        )
        return positions[valid], rotations[valid], scores[valid], details[valid]


class PeakCallerSort(PeakCaller):
    """The ``num_peaks`` highest scores of the volume (:644-659)."""

    def call_peaks(self, scores, **kwargs):
        t = _torch()
        flat = scores.reshape(-1)
        k = min(self.num_peaks, be.size(flat))
        coords = be.topk_indices(scores, k)
        return t.stack(coords, dim=1), None


class PeakCallerMaximumFilter(PeakCaller):
    """Local maxima under a box maximum filter of edge ``min_distance``,
    then the distance constraint (:662-670)."""

    def call_peaks(self, scores, **kwargs):
        return be.max_filter_coordinates(scores, self.min_distance), None
