"""Template matching data container (mirrors ``tme/matching_data.py``).

Holds target, template, masks, filters and rotations; computes the Fourier
padding and shift of the scan; extracts reflect-padded target subsets for
split scans.  Batch dimensions (stacks of targets/templates) are not part of
this build: every array is one d-dimensional volume, so all batch masks are
zero.
"""

import warnings
from typing import List, Optional, Tuple

import numpy as np

from .density import Density
from .backends import backend as be

__all__ = ["MatchingData"]


class MatchingData:
    """Data required for template matching.

    Parameters
    ----------
    target, template : array
        Target and template volumes (same dimensionality).
    template_mask, target_mask : array, optional
        Masks; the template mask defaults to all ones.
    invert_target : bool, optional
        Negate the target before matching.
    rotations : array, optional
        (d, d) or (n, d, d) rotation matrices; identity by default.
    """

    def __init__(
        self,
        target,
        template,
        template_mask=None,
        target_mask=None,
        invert_target: bool = False,
        rotations=None,
    ):
        self._target = target
        self._target_mask = None
        self.target_mask = target_mask
        self._template_mask = None
        self.template = template
        if template_mask is not None:
            self.template_mask = template_mask
        self.rotations = rotations
        self._invert_target = invert_target
        self._translation_offset = tuple(0 for _ in range(len(target.shape)))
        self._template_filter = None
        self._target_filter = None
        self._set_matching_dimension()

    # ------------------------------------------------------------ dimensions
    def set_matching_dimension(self, target_dim: int = None, template_dim: int = None):
        if target_dim not in (None, ()) or template_dim not in (None, ()):
            raise NotImplementedError("batched targets/templates are not supported by this build")
        self._set_matching_dimension()

    def _set_matching_dimension(self, target_dims: Tuple[int] = (), template_dims: Tuple[int] = ()):
        self._target_dim, self._template_dim = tuple(target_dims), tuple(template_dims)
        ndim = len(self._target.shape)
        if len(self._template.shape) != ndim:
            raise ValueError("target and template must have the same number of dimensions")
        self._output_target_shape = tuple(int(x) for x in self._target.shape)
        self._output_template_shape = tuple(int(x) for x in self._template.shape)
        self._batch_mask = tuple(0 for _ in range(ndim))
        self._template_batch = tuple(0 for _ in range(ndim))
        self._target_batch = tuple(0 for _ in range(ndim))
        self._output_shape = self._output_target_shape

    @staticmethod
    def _batch_shape(shape: Tuple[int], mask: Tuple[int], keepdims=True) -> Tuple[int]:
        if keepdims:
            return tuple(int(x) if y == 0 else 1 for x, y in zip(shape, mask))
        return tuple(int(x) for x, y in zip(shape, mask) if y == 0)

    @staticmethod
    def _batch_axis(mask: Tuple[int]) -> Tuple[int]:
        return tuple(i for i in range(len(mask)) if mask[i] == 0)

    @staticmethod
    def _batch_iter(shape: Tuple[int], mask: Tuple[int]):
        yield tuple(slice(None) for _ in shape)

    # ---------------------------------------------------------------- subsets
    @staticmethod
    def _shape_to_slice(shape: Tuple[int]) -> Tuple[slice]:
        return tuple(slice(0, int(dim)) for dim in shape)

    def subset_array(self, arr, arr_slice: Tuple[slice], padding, invert: bool = False) -> np.ndarray:
        """Extract ``arr[arr_slice]`` grown by ``padding`` (split scans).

        The padding is split ``left = ceil(p/2)``, ``right = floor(p/2)`` and
        taken from the array where available; what falls outside the array
        is filled by reflection (tme/matching_data.py:96-169).
        """
        padding = np.maximum(np.asarray(be.to_numpy_array(padding)), 0).astype(int)
        start = np.array([s.start for s in arr_slice], dtype=int)
        stop = np.array([s.stop for s in arr_slice], dtype=int)
        right_pad = np.floor_divide(padding, 2)
        left_pad = right_pad + np.mod(padding, 2)
        take_left = np.minimum(start, left_pad)
        take_right = np.minimum(np.subtract(arr.shape, stop), right_pad).astype(int)
        box = tuple(slice(int(a), int(b)) for a, b in zip(start - take_left, stop + take_right))
        if isinstance(arr, Density):
            # a memory-mapped map is re-read for just this box (row reads)
            if isinstance(arr.data, np.memmap):
                sub = Density.from_file(arr.data.filename, subset=box).data
            else:
                sub = np.asarray(arr.data[box])
        else:
            sub = np.asarray(be.to_numpy_array(arr[box]) if not isinstance(arr, np.ndarray) else arr[box])
        pad = tuple((int(l), int(r)) for l, r in zip(left_pad - take_left, right_pad - take_right))
        if any(l or r for l, r in pad):
            sub = np.pad(sub, pad, mode="reflect")
        else:
            sub = np.array(sub, copy=True)
        return -sub if invert else sub

    def subset_by_slice(
        self,
        target_slice: Tuple[slice] = None,
        template_slice: Tuple[slice] = None,
        target_pad=None,
        template_pad=None,
        invert_target: bool = False,
    ) -> Tuple["MatchingData", Tuple]:
        """New instance restricted to the given slices, plus its offset."""
        if target_slice is None:
            target_slice = self._shape_to_slice(self._target.shape)
        if template_slice is None:
            template_slice = self._shape_to_slice(self._template.shape)
        if target_pad is None:
            target_pad = np.zeros(len(self._target.shape), dtype=int)
        if template_pad is None:
            template_pad = np.zeros(len(self._template.shape), dtype=int)

        target = self.subset_array(self._target, target_slice, target_pad, invert=self._invert_target)
        template = self.subset_array(self._template, template_slice, template_pad)
        target_mask = template_mask = None
        if self._target_mask is not None:
            sl = tuple(s if n != 1 else slice(0, 1) for s, n in zip(target_slice, self._target_mask.shape))
            target_mask = self.subset_array(self._target_mask, sl, target_pad)
        if self._template_mask is not None:
            sl = tuple(s if n != 1 else slice(0, 1) for s, n in zip(template_slice, self._template_mask.shape))
            template_mask = self.subset_array(self._template_mask, sl, template_pad)

        ret = self.__class__(
            target=target,
            template=template,
            template_mask=template_mask,
            target_mask=target_mask,
            rotations=self.rotations,
            invert_target=self._invert_target,
        )
        offset = tuple(int(s.start) for s in target_slice)
        ret.target_filter = self.target_filter
        ret.template_filter = self.template_filter
        return ret, offset

    def to_backend(self):
        """Move arrays to the active backend; volumes become float dtype."""
        t = type(be.zeros((1), dtype=be._float_dtype))
        for name, value in list(vars(self).items()):
            if isinstance(value, np.ndarray):
                conv = be.to_backend_array(value.copy())
            elif isinstance(value, t):
                conv = be.to_backend_array(value)
            else:
                continue
            current = be.get_fundamental_dtype(conv)
            target_dtype = be._fundamental_dtypes[current]
            if name in ("_template", "_template_mask", "_target", "_target_mask"):
                target_dtype = be._float_dtype
            if conv.dtype != target_dtype:
                conv = be.astype(conv, target_dtype)
            setattr(self, name, conv)

    # --------------------------------------------------------------- padding
    def target_padding(self, pad_target: bool = False) -> Tuple[int]:
        """Target padding to the full convolution shape (template - 1)."""
        if pad_target:
            return tuple(int(x) - 1 for x in self._output_template_shape)
        return tuple(0 for _ in self._output_target_shape)

    @staticmethod
    def _fourier_padding(
        target_shape: Tuple[int],
        template_shape: Tuple[int],
        pad_target: bool = False,
        batch_mask: Tuple[int] = None,
    ) -> Tuple[Tuple, Tuple, Tuple, Tuple]:
        """Transform shapes and the roll aligning scores with the target.

        conv = max(target, template); fast = next fast length of conv;
        fast_ft = fast[:-1] + (fast[-1]//2 + 1);
        shift = 1 - T//2 - T%2 (one less per axis when the target was padded
        and T is even) (tme/matching_data.py:489-532).
        """
        target_shape = np.asarray(target_shape, dtype=int)
        template_shape = np.asarray(template_shape, dtype=int)
        pad_shape = np.maximum(target_shape, template_shape)
        conv, fast, fast_ft = be.compute_convolution_shapes(pad_shape, np.ones_like(pad_shape))
        template_mod = np.mod(template_shape, 2)
        shift = 1 - np.floor_divide(template_shape, 2) - template_mod
        diff = target_shape - template_shape
        if np.any(diff < 0):
            warnings.warn(
                "Template is larger than target and padding is turned off. Consider "
                "swapping them or activate padding. Correcting the shift for now."
            )
            neg = diff < 0
            shift = np.subtract(shift, (np.divide(diff, 2) + np.mod(diff, 2)) * neg).astype(int)
        if pad_target:
            shift = shift - (1 - template_mod)
        return (
            tuple(int(x) for x in conv),
            tuple(int(x) for x in fast),
            tuple(int(x) for x in fast_ft),
            tuple(int(x) for x in shift),
        )

    def fourier_padding(self, pad_target: bool = False) -> Tuple:
        return self._fourier_padding(
            target_shape=self._output_target_shape,
            template_shape=self._output_template_shape,
            pad_target=pad_target,
        )

    # -------------------------------------------------------------- rotations
    @property
    def rotations(self):
        return self._rotations

    @rotations.setter
    def rotations(self, rotations):
        if rotations is None:
            rotations = np.eye(len(self._target.shape))
        rotations = np.asarray(be.to_numpy_array(rotations))
        if rotations.ndim not in (2, 3):
            raise ValueError("Rotations have to be a rank 2 or 3 array.")
        if rotations.ndim == 2:
            rotations = rotations.reshape(1, *rotations.shape)
        self._rotations = rotations.astype(np.float32)

    # ---------------------------------------------------------------- arrays
    @staticmethod
    def _reversed(arr):
        if arr is None:
            return None
        return be.reverse(arr) if not isinstance(arr, np.ndarray) else arr[(slice(None, None, -1),) * arr.ndim]

    @property
    def target(self):
        return self._target.data if isinstance(self._target, Density) else self._target

    @target.setter
    def target(self, arr):
        self._target = arr

    @property
    def target_mask(self):
        return self._target_mask

    @target_mask.setter
    def target_mask(self, arr):
        self._target_mask = self._set_mask(arr, self._target.shape)

    @property
    def template(self):
        """The template reversed along every axis (correlation without conjugation)."""
        return self._reversed(self._template)

    @template.setter
    def template(self, arr):
        self._template = arr.data if isinstance(arr, Density) else arr
        if getattr(self, "_template_mask", None) is None:
            self._template_mask = np.ones(arr.shape, dtype=np.float32)

    @property
    def template_mask(self):
        """The template mask reversed along every axis."""
        return self._reversed(getattr(self, "_template_mask", None))

    @template_mask.setter
    def template_mask(self, arr):
        self._template_mask = self._set_mask(arr, self._template.shape)

    @staticmethod
    def _set_mask(mask, shape):
        if isinstance(mask, Density):
            mask = mask.data
        if mask is not None and np.broadcast_shapes(tuple(mask.shape), tuple(shape)) != tuple(shape):
            raise ValueError("Mask and data shape need to be broadcastable.")
        return mask

    @property
    def template_filter(self):
        return self._template_filter

    @template_filter.setter
    def template_filter(self, value):
        self._template_filter = value

    @property
    def target_filter(self):
        return self._target_filter

    @target_filter.setter
    def target_filter(self, value):
        self._target_filter = value

    def computation_schedule(self, matching_method: str = "FLCSphericalMask", max_cores: int = 1,
                             use_gpu: bool = False, pad_fourier: bool = False, pad_target_edges: bool = False,
                             analyzer_method: str = None, available_memory: int = None, max_splits: int = 256):
        """Target splits and (outer, inner) jobs that fit ``available_memory``
        (tme/matching_data.py:559-634); on the hip backend the budget is the
        device's free HBM and the estimate includes the fused scan's batch
        workspace.  GPU schedules split only into outer jobs."""
        from .backends import backend as be
        from .matching_utils import compute_parallelization_schedule

        if available_memory is None:
            available_memory = be.get_available_memory() * be.device_count()
        shape1 = self._output_target_shape
        shape2 = tuple(0 for _ in self._output_template_shape)
        if pad_fourier:
            shape2 = self._output_template_shape
        padding = tuple(0 for _ in shape1)
        if pad_target_edges:
            padding = tuple(int(x) for x in self._output_template_shape)
        dt = getattr(be, "datatype_bytes", None)
        nbytes = (4, 8, 4) if dt is None else (dt(be._float_dtype), dt(be._complex_dtype), dt(be._int_dtype))
        return compute_parallelization_schedule(
            shape1=shape1, shape2=shape2, shape1_padding=padding, max_cores=max_cores, max_ram=available_memory,
            matching_method=matching_method, analyzer_method=analyzer_method,
            backend=getattr(be, "_backend_name", None), float_nbytes=nbytes[0], complex_nbytes=nbytes[1],
            integer_nbytes=nbytes[2], split_only_outer=use_gpu,
            split_axes=self._target_dim if len(self._target_dim) else None, max_splits=max_splits,
        )

    def _split_rotations_on_jobs(self, n_jobs: int) -> List[np.ndarray]:
        """Contiguous rotation chunks; the last chunk takes the remainder."""
        per = self.rotations.shape[0] // n_jobs
        out = []
        for n in range(n_jobs):
            end = None if n == n_jobs - 1 else (n + 1) * per
            out.append(self.rotations[n * per : end])
        return out

    def _free_data(self):
        for name in ("_target", "_template", "_template_mask", "_target_mask"):
            setattr(self, name, None)
