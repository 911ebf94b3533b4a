"""Maximum score over rotations (mirrors ``tme/analyzer/aggregation.py:31-431``)."""

import os
import tempfile
from typing import Dict, List, Tuple

import numpy as np

from ..backends import backend as be
from ..matching_utils import apply_convolution_mode
from .base import AbstractAnalyzer

__all__ = ["MaxScoreOverRotations", "assign_rotation_indices"]


def _to_memmap(arr: np.ndarray) -> np.ndarray:
    fd, name = tempfile.mkstemp(suffix=".npy")
    os.close(fd)
    mm = np.memmap(name, mode="w+", shape=arr.shape, dtype=arr.dtype)
    mm[:] = arr
    mm.flush()
    return np.memmap(name, mode="r", shape=arr.shape, dtype=arr.dtype)


def assign_rotation_indices(mapping: Dict, rotation_matrices, inversion_mapping: bool) -> List[int]:
    """Rotation indices of a batch of rotations as successive analyzer calls
    assign them, updating ``mapping`` the same way
    (tme/analyzer/aggregation.py:170-189): with ``inversion_mapping`` the
    index is the call number and the mapping stores index -> matrix;
    otherwise matrices are keyed by their float bytes and a repeated matrix
    reuses its index."""
    out = []
    for i in range(rotation_matrices.shape[0]):
        rot = be.astype(rotation_matrices[i], be._float_dtype)
        idx = len(mapping)
        if inversion_mapping:
            mapping[idx] = rot
        else:
            idx = mapping.setdefault(be.tobytes(rot), idx)
        out.append(idx)
    return out


class MaxScoreOverRotations(AbstractAnalyzer):
    """Best score of every translation over all rotations, and its rotation.

    State: ``(scores, rotations, rotation_mapping)``.  ``scores`` starts at
    ``score_threshold`` and ``rotations`` at -1; an update replaces a voxel
    only where the new score is strictly greater, so the earliest rotation
    wins exact ties (npfftw_backend.py:503-535).

    With ``inversion_mapping`` (set by ``scan`` when one job owns the
    analyzer) the rotation index is the call order and the mapping stores
    index -> matrix; otherwise rotations are keyed by their bytes so repeated
    matrices share an index.
    """

    def __init__(
        self,
        shape: Tuple[int],
        offset=None,
        score_threshold: float = 0,
        shm_handler: object = None,
        use_memmap: bool = False,
        inversion_mapping: bool = False,
        jax_mode: bool = False,
        **kwargs,
    ):
        self._use_memmap = use_memmap
        self._score_threshold = score_threshold
        self._shape = tuple(int(x) for x in shape)
        self._inversion_mapping = inversion_mapping and not jax_mode
        self._jax_mode = jax_mode
        if offset is None:
            offset = np.zeros(len(self._shape), dtype=int)
        self._offset = np.asarray(be.to_numpy_array(offset), dtype=int)

    @property
    def shareable(self) -> bool:
        return True

    def init_state(self):
        scores = be.full(self._shape, fill_value=self._score_threshold, dtype=be._float_dtype)
        rotations = be.full(self._shape, fill_value=-1, dtype=be._int_dtype)
        return scores, rotations, {}

    def rotation_indices(self, state, rotation_matrices) -> List[int]:
        """Assign indices to a batch of rotations exactly as successive calls would."""
        return assign_rotation_indices(state[2], rotation_matrices, self._inversion_mapping)

    def __call__(self, state, scores, rotation_matrix, **kwargs):
        prev_scores, rotations, mapping = state
        if self._jax_mode:
            idx = kwargs.get("rotation_index", 0)
        else:
            idx = self.rotation_indices(state, rotation_matrix[None])[0]
        scores, rotations = be.max_score_over_rotations(
            scores=scores, max_scores=prev_scores, rotations=rotations, rotation_index=idx
        )
        return scores, rotations, mapping

    @staticmethod
    def _invert_rmap(rotation_mapping: Dict) -> Dict:
        """bytes -> index mapping to index -> float32 matrix."""
        out, ndim = {}, None
        for k, v in rotation_mapping.items():
            rmat = np.frombuffer(k, dtype=np.float32)
            if ndim is None:
                ndim = int(round(np.sqrt(rmat.size)))
            out[v] = rmat.reshape(ndim, ndim)
        return out

    def result(
        self,
        state,
        targetshape: Tuple[int] = None,
        templateshape: Tuple[int] = None,
        convolution_shape: Tuple[int] = None,
        fourier_shift: Tuple[int] = None,
        convolution_mode: str = None,
        **kwargs,
    ) -> Tuple:
        """Roll by ``fourier_shift``, crop by ``convolution_mode``, move to host."""
        scores, rotations, mapping = state
        if fourier_shift is not None:
            axis = tuple(range(len(fourier_shift)))
            scores = be.roll(scores, shift=fourier_shift, axis=axis)
            rotations = be.roll(rotations, shift=fourier_shift, axis=axis)
        if convolution_mode is not None:
            kw = dict(
                s1=targetshape, s2=templateshape, convolution_mode=convolution_mode,
                convolution_shape=convolution_shape,
            )
            scores = apply_convolution_mode(scores, **kw)
            rotations = apply_convolution_mode(rotations, **kw)
        scores = np.ascontiguousarray(be.to_numpy_array(scores))
        rotations = np.ascontiguousarray(be.to_numpy_array(rotations))
        if self._use_memmap:
            scores, rotations = _to_memmap(scores), _to_memmap(rotations)
        if self._inversion_mapping:
            keys = list(mapping.keys())
            if keys:
                mats = [mapping[k] for k in keys]
                stacked = np.asarray(
                    be.to_numpy_array(be._array_backend.stack(mats) if not isinstance(mats[0], np.ndarray)
                                      else np.stack(mats)),
                    dtype=np.float32,
                )
                mapping = {stacked[i].tobytes(): k for i, k in enumerate(keys)}
            else:
                mapping = {}
        return scores, np.asarray(self._offset), rotations, self._invert_rmap(mapping)

    @staticmethod
    def _harmonize_states(states: List[Tuple]):
        master, out_shape = {}, None
        for i, st in enumerate(states):
            if st is None:
                continue
            scores, offset, rotations, mapping = st
            if out_shape is None:
                out_shape = np.zeros(scores.ndim, int)
            out_shape = np.maximum(out_shape, np.add(offset, scores.shape))
            local = {}
            for key, value in mapping.items():
                b = np.asarray(value, dtype=np.float32).tobytes()
                local[b] = key
                if b not in master:
                    master[b] = len(master)
            states[i] = (scores, offset, rotations, local)
        if out_shape is not None:
            out_shape = tuple(int(x) for x in out_shape)
        return master, out_shape, states

    @classmethod
    def merge(cls, results: List[Tuple], **kwargs) -> Tuple:
        """Combine split results by offset: strict '>' max, rotation re-indexing."""
        use_memmap = kwargs.get("use_memmap", False)
        results = list(results)
        if len(results) == 1:
            ret = results[0]
            if use_memmap:
                s, o, r, m = ret
                ret = (_to_memmap(np.asarray(s)), o, _to_memmap(np.asarray(r)), m)
            return ret
        master, out_shape, results = cls._harmonize_states(results)
        if out_shape is None:
            return None
        sdtype = results[0][0].dtype
        rdtype = results[0][2].dtype
        scores_out = np.full(out_shape, fill_value=kwargs.get("score_threshold", 0), dtype=sdtype)
        rot_out = np.full(out_shape, fill_value=-1, dtype=rdtype)
        for i in range(len(results)):
            if results[i] is None:
                continue
            scores, offset, rotations, mapping = results[i]
            box = tuple(slice(int(a), int(a) + int(n)) for a, n in zip(offset, scores.shape))
            upd = scores > scores_out[box]
            scores_out[box][upd] = scores[upd]
            lut = np.arange(len(mapping) + 1, dtype=rot_out.dtype)
            for key, value in mapping.items():
                lut[value] = master[key]
            sel = rotations[upd]
            if len(sel):
                rot_out[box][upd] = lut[sel]
            results[i] = None
        if use_memmap:
            scores_out, rot_out = _to_memmap(scores_out), _to_memmap(rot_out)
        return scores_out, np.zeros(scores_out.ndim, dtype=int), rot_out, cls._invert_rmap(master)
