"""Abstract matching backend: the array/FFT strategy the exhaustive scan uses.

Mirrors the interface of ``tme/backends/matching_backend.py:28-1271``.  Only
the methods the exhaustive-scan path calls are abstract here (the list in
the reference's hot path: conversion, allocation, shape, arithmetic,
numeric info, layout, sharing, device, FFT, rotation and scoring).  Attribute
access that the backend does not define falls through to the wrapped array
module, as in the reference (``matching_backend.py:108-127``).
"""

from abc import ABC, abstractmethod
from typing import Any, Callable, Dict, List, Tuple


def _create_metafunction(func_name: str) -> Callable:
    """Return a method forwarding to ``self._array_backend.<func_name>``."""

    def metafunction(self, *args, **kwargs) -> Any:
        return getattr(self._array_backend, func_name)(*args, **kwargs)

    metafunction.__name__ = func_name
    return metafunction


class MatchingBackend(ABC):
    """Strategy class for template matching backends.

    Parameters
    ----------
    array_backend : module
        Array module that provides the non-specialised operations.
    float_dtype, complex_dtype, int_dtype, overflow_safe_dtype : dtype
        Dtypes of real, complex, integer and reduction-safe arrays.
    """

    def __init__(self, array_backend, float_dtype, complex_dtype, int_dtype, overflow_safe_dtype):
        self._array_backend = array_backend
        self._float_dtype = float_dtype
        self._complex_dtype = complex_dtype
        self._int_dtype = int_dtype
        self._overflow_safe_dtype = overflow_safe_dtype
        self._fundamental_dtypes = {
            int: self._int_dtype,
            float: self._float_dtype,
            complex: self._complex_dtype,
        }

    def __getattr__(self, name: str):
        if name.startswith("__") or name == "_array_backend":
            raise AttributeError(name)
        return getattr(self._array_backend, name)

    def __dir__(self) -> List:
        attrs = set(dir(self.__class__)) | set(self.__dict__.keys())
        attrs |= set(dir(self._array_backend))
        return sorted(attrs)

    # conversion -----------------------------------------------------------
    @abstractmethod
    def to_backend_array(self, arr):
        """Convert ``arr`` to a backend array."""

    @abstractmethod
    def to_numpy_array(self, arr):
        """Convert a backend array to :obj:`numpy.ndarray`."""

    @abstractmethod
    def to_cpu_array(self, arr):
        """Move a backend array to host memory (backend type)."""

    @abstractmethod
    def get_fundamental_dtype(self, arr) -> type:
        """Return ``int``, ``float`` or ``complex`` for ``arr``'s dtype."""

    @abstractmethod
    def free_cache(self):
        """Release cached allocations."""

    # arithmetic -----------------------------------------------------------
    @abstractmethod
    def add(self, x1, x2, out=None):
        """Elementwise ``x1 + x2``."""

    @abstractmethod
    def subtract(self, x1, x2, out=None):
        """Elementwise ``x1 - x2``."""

    @abstractmethod
    def multiply(self, x1, x2, out=None):
        """Elementwise ``x1 * x2``."""

    @abstractmethod
    def divide(self, x1, x2, out=None):
        """Elementwise ``x1 / x2``."""

    @abstractmethod
    def sum(self, arr, axis=None, keepdims: bool = False):
        """Sum over ``axis``."""

    @abstractmethod
    def maximum(self, x1, x2, out=None):
        """Elementwise maximum."""

    @abstractmethod
    def sqrt(self, arr, out=None):
        """Elementwise square root."""

    @abstractmethod
    def square(self, arr, out=None):
        """Elementwise square."""

    # allocation / info ----------------------------------------------------
    @abstractmethod
    def size(self, arr) -> int:
        """Number of elements."""

    @abstractmethod
    def fill(self, arr, value):
        """Fill ``arr`` with ``value`` and return it."""

    @abstractmethod
    def zeros(self, shape, dtype):
        """Zero-initialised array."""

    @abstractmethod
    def full(self, shape, fill_value, dtype=None):
        """Array filled with ``fill_value``."""

    @abstractmethod
    def eps(self, dtype):
        """Machine epsilon of ``dtype``."""

    @abstractmethod
    def datatype_bytes(self, dtype) -> int:
        """Bytes per element of ``dtype``."""

    @abstractmethod
    def astype(self, arr, dtype):
        """Cast ``arr`` to ``dtype``."""

    def tobytes(self, arr) -> bytes:
        """Byte string of ``arr`` (used to key rotations)."""
        return self.to_numpy_array(arr).tobytes()

    # layout ---------------------------------------------------------------
    @abstractmethod
    def roll(self, a, shift, axis):
        """Circular shift."""

    @abstractmethod
    def topleft_pad(self, arr, shape: Tuple[int], padval: float = 0):
        """Place ``arr`` in the top-left corner of a new array of ``shape``."""

    @abstractmethod
    def reverse(self, arr, axis: Tuple[int] = None):
        """Reverse ``arr`` along ``axis`` (all axes by default)."""

    # sharing --------------------------------------------------------------
    @abstractmethod
    def from_sharedarr(self, args):
        """Inverse of :meth:`to_sharedarr`."""

    @abstractmethod
    def to_sharedarr(self, arr, shared_memory_handler=None):
        """Make ``arr`` shareable across processes."""

    # device ---------------------------------------------------------------
    @abstractmethod
    def get_available_memory(self) -> int:
        """Free memory in bytes on the active device."""

    @abstractmethod
    def set_device(self, device_index: int):
        """Context manager selecting the active device."""

    @abstractmethod
    def device_count(self) -> int:
        """Number of devices."""

    # FFT ------------------------------------------------------------------
    @abstractmethod
    def build_fft(
        self,
        fwd_shape: Tuple[int],
        inv_shape: Tuple[int],
        real_dtype,
        cmpl_dtype,
        inv_output_shape: Tuple[int] = None,
        temp_fwd=None,
        temp_inv=None,
        fwd_axes: Tuple[int] = None,
        inv_axes: Tuple[int] = None,
        **kwargs,
    ) -> Tuple[Callable, Callable]:
        """Return ``(rfftn(arr, out) -> out, irfftn(arr, out) -> out)``."""

    @abstractmethod
    def compute_convolution_shapes(
        self, arr1_shape: Tuple[int], arr2_shape: Tuple[int]
    ) -> Tuple[List[int], List[int], List[int]]:
        """Return ``(convolution_shape, fast_shape, fast_ft_shape)``."""

    # rotation -------------------------------------------------------------
    @abstractmethod
    def rigid_transform(
        self,
        arr,
        rotation_matrix,
        arr_mask=None,
        translation=None,
        use_geometric_center: bool = False,
        out=None,
        out_mask=None,
        order: int = 3,
        cache: bool = False,
        batched: bool = False,
    ):
        """Rotate ``arr`` (and ``arr_mask``) by ``rotation_matrix``."""

    # scoring --------------------------------------------------------------
    @abstractmethod
    def norm_scores(self, arr, exp_sq, sq_exp, n_obs, eps, out):
        """Normalise ``arr`` by the local standard deviation."""

    @abstractmethod
    def max_score_over_rotations(self, scores, max_scores, rotations, rotation_index: int):
        """Fold ``scores`` into the running maximum and its rotation index."""

    def extract_center(self, arr, newshape: Tuple[int]):
        """Central ``newshape`` box of ``arr``."""
        starts = [(int(c) - int(n)) // 2 for c, n in zip(arr.shape, newshape)]
        box = tuple(slice(s, s + int(n)) for s, n in zip(starts, newshape))
        return arr[box]

    @property
    def backend_args(self) -> Dict:
        return getattr(self, "_init_kwargs", {})
